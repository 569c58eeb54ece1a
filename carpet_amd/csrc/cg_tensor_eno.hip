// cg_tensor_eno.hip -- per-call tensor launches for LK_ENO line operators.
#include "cg_tensor_plain.hpp"

namespace cg {
template int launch_tensor_kind<LK_ENO>(int, int, int, const TensorArgs &, const TileCfg &,
                                         hipStream_t);
} // namespace cg
