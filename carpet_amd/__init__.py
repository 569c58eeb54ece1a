"""carpet_amd: CarpetLib's inter-grid transfer operators on MI355X (gfx950).

The operators run as hand-written HIP kernels in libcarpet_amd.so behind a C
ABI (include/carpet_amd.h); this package is the host-side mirror of the
reference interface used by tests, benchmarks and Python callers.
"""
from ._lib import CarpetError, lib  # noqa: F401  (fails loudly if the library is missing)
from .core import (  # noqa: F401
    CELL_CENTERED,
    OP_COPY,
    OP_ENO,
    OP_WENO,
    OP_LAGRANGE,
    OP_NONE,
    OP_RESTRICT,
    OP_SYNC,
    OP_TVD,
    PREC_F64,
    PREC_MIXED,
    PREC_MIXED_DOT2,
    VERTEX_CENTERED,
    Batch,
    Box,
    DeviceArray,
    Event,
    Stream,
    copy_3d,
    interpolate_3d_2tl,
    interpolate_3d_3tl,
    interpolate_eno_3d_3tl,
    pad_shape,
    prolongate_3d_cc_eno_rf2,
    prolongate_3d_eno,
    prolongate_3d_cc_rf2,
    prolongate_3d_cc_tvd,
    prolongate_3d_rf2,
    prolongate_3d_rf2_mixed,
    prolongate_3d_cc_rf2_mixed,
    relative_errors,
    transfer_from_mixed,
    restrict_3d_cc_o3_rf2,
    restrict_3d_cc_o5_rf2,
    restrict_3d_cc_rf2,
    restrict_3d_rf2,
    set_restriction,
    synchronize,
    transfer_from,
)
