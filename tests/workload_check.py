"""Replays a carpet_amd.workloads.Workload on the CPU oracle.

The dispatch below restates data<T>::transfer_time/transfer_p_r
(CarpetLib/src/data.cc:485-556) on top of the oracle's operators, so a
workload's GPU outputs can be compared array by array.
"""
from __future__ import annotations

import numpy as np

import pyoracle as orc

OP_COPY, OP_LAGRANGE, OP_ENO, OP_WENO, OP_TVD, OP_LAGRANGE_MONOTONE = 4, 5, 6, 7, 8, 9
VERTEX, CELL = 1, 2


def _host(spec, init):
    lo, up, st = spec.box.lower, spec.box.upper, spec.box.stride
    return orc.HostArr((lo, up, st), np.ascontiguousarray(init, dtype=np.float64).copy())


def _b(box):
    return (box.lower, box.upper, box.stride)


def host_state(w) -> dict:
    """Oracle arrays holding every array's initial extent."""
    inputs = w.host_inputs()
    return {k: _host(s, inputs[k]) for k, s in w.arrays.items()}


def run_oracle(w, host=None, steps=None) -> dict:
    """Replays w's transfers (or `steps`) in phase order on the oracle,
    starting from `host` (updated in place) or from the initial arrays;
    returns the outputs."""
    if host is None:
        host = host_state(w)
    for t in sorted(w.steps if steps is None else steps, key=lambda t: t.phase):
        dst = host[t.dst]
        srcs = [host[s] for s in t.srcs]
        sst, dstst = srcs[0].box[2], dst.box[2]
        rc, tl0, ntl = orc.find_source_timelevel(t.times, t.time, t.order_time, t.op)
        assert rc == 0, (t.name, rc)
        reg, sreg = _b(t.dstreg), _b(t.srcreg)
        if sst == dstst:
            if ntl == 1:
                rc = orc.copy_3d(srcs[tl0], dst, sreg, reg, t.slab)
            else:
                rc = orc.interpolate_3d(srcs[:ntl], t.times[:ntl], dst, t.time, reg)
        elif all(a > b for a, b in zip(sst, dstst)):
            if t.op in (OP_COPY, OP_LAGRANGE):
                kind, order = (orc.KIND_VLAG if t.cent == VERTEX else orc.KIND_CLAG), t.order_space
            elif t.op == OP_LAGRANGE_MONOTONE:
                kind, order = orc.KIND_MONO5, 5
            elif t.op in (OP_ENO, OP_WENO) and t.cent == VERTEX:
                kind, order = orc.KIND_VENO, 3
            elif t.op == OP_ENO:
                kind, order = orc.KIND_ENO, (3 if t.order_space == 5 else 2)
            else:
                kind, order = orc.KIND_TVD, 1
            rc = orc.prolongate_time(kind, order, srcs, t.times, dst, t.time, reg, t.order_time)
        else:
            f = orc.restrict_3d_rf2 if t.cent == VERTEX else orc.restrict_3d_cc_rf2
            rc = f(srcs[tl0], dst, reg)
        assert rc == 0, (t.name, rc)
    return {k: host[k].data for k in w.outputs if k in host}
