"""Pins the oracle's cell-centred ENO stencil choice
(prolongate_3d_cc_eno_rf2.cc:231-408) with data whose smoothest stencil is
known by construction.

Every line is a polynomial of degree ORDER with one or two kinks: a term
A*(x - t)_+^(ORDER-1) (or its mirror) added on one side of a point t placed
between two coarse values.  The ORDER-th undivided difference of such a term
is a B-spline value: zero for a stencil that does not straddle t, positive
for one that does.  So every candidate stencil keeps the sign of the
polynomial's difference (no first-order fallback) and the stencils that avoid
the kinks are strictly the smoothest.  The expected value is that stencil's
Lagrange interpolant at the fine point, computed in exact rationals from the
node positions (no coefficient table involved), and every alternative
stencil gives a different value, so choosing the roughest stencil, swapping
left and right, or breaking ties the other way fails these tests.
"""
from fractions import Fraction as F

import numpy as np
import pytest

import pyoracle as orc
from helpers import host_array, sizes


def lagrange(nodes, vals, x0):
    """Value at x0 of the polynomial through (nodes, vals), in rationals."""
    r = F(0)
    for i, (xi, vi) in enumerate(zip(nodes, vals)):
        w = F(1)
        for j, xj in enumerate(nodes):
            if j != i:
                w *= (x0 - xj) / (xi - xj)
        r += w * vi
    return r


def undivided(vals):
    """Highest-order undivided difference of consecutive values."""
    v = list(vals)
    for _ in range(len(vals) - 1):
        v = [b - a for a, b in zip(v, v[1:])]
    return v[0]


def window(order, di):
    """Coarse offsets (relative to the fine point's base cell) the 1-D
    operator reads, and their positions: coarse value i sits at i + 1/2, the
    fine point at 1/4 + di/2 (prolongate_3d_cc_eno_rf2.cc:17-40)."""
    rc, mm, mx = orc.eno_window(order, di)
    # ENO<2> chooses between 2 stencils of 3 points, ENO<3> among 3 of 4
    assert rc == 0 and mx - mm == {2: 4, 3: 6}[order]
    return mm, [F(2 * i + 1, 2) for i in range(mm, mx)], F(1, 4) + F(di, 2)


def kinked_line(order, di, target, amp=4):
    """Values on the window: (x+4)^ORDER plus kinks just outside the nodes of
    candidate stencil `target` (0 = leftmost).  Returns (values, candidate
    node lists, positions, fine point, window start)."""
    mm, pos, x0 = window(order, di)
    n, ns = len(pos), order + 1
    cands = [list(range(s, s + ns)) for s in range(n - ns + 1)]
    f = [(x + 4) ** order for x in pos]
    tgt, m = cands[target], order - 1
    if tgt[0] > 0:  # kink left of the target's first node
        t = (pos[tgt[0] - 1] + pos[tgt[0]]) / 2
        # (t-x)_+^m has the sign (-1)^ORDER in its ORDER-th difference
        f = [v + amp * (-1) ** order * (t - x) ** m if x < t else v for v, x in zip(f, pos)]
    if tgt[-1] < n - 1:  # kink right of the target's last node
        t = (pos[tgt[-1]] + pos[tgt[-1] + 1]) / 2
        f = [v + amp * (x - t) ** m if x > t else v for v, x in zip(f, pos)]
    return f, cands, pos, x0, mm


def _check_construction(f, cands, target, mm, di, pos, x0):
    """The data has what the pin claims: one sign of every candidate's
    undivided difference, a strict minimum at `target`, a result between the
    two nearest coarse values (so the final monotonicity check keeps it), and
    a different value from every other candidate."""
    V = [undivided([f[k] for k in c]) for c in cands]
    assert all(v > 0 for v in V) or all(v < 0 for v in V), V
    assert all(abs(V[target]) < abs(v) for i, v in enumerate(V) if i != target), V
    vals = [lagrange([pos[k] for k in c], [f[k] for k in c], x0) for c in cands]
    a, b = f[-mm - 1 + di], f[-mm + di]
    assert min(a, b) < vals[target] < max(a, b)
    assert len(set(vals)) == len(vals), vals
    return vals


@pytest.mark.parametrize("order,target", [(2, 0), (2, 1), (3, 0), (3, 1), (3, 2)])
@pytest.mark.parametrize("di", [0, 1])
def test_eno_line_picks_smoothest_stencil(order, di, target):
    f, cands, pos, x0, mm = kinked_line(order, di, target)
    vals = _check_construction(f, cands, target, mm, di, pos, x0)
    got = orc.eno1(order, di, np.array([float(v) for v in f]))
    # dyadic data and weights: the double result is exact
    assert F(got) == vals[target], (got, [float(v) for v in vals])


@pytest.mark.parametrize("di", [0, 1])
def test_eno3_tie_between_left_and_right_goes_left(di):
    """ENO<3> keeps the first strict minimum of |V| starting from the centred
    stencil, scanning left, centred, right (:345-349): with |V_left| =
    |V_right| < |V_centred| the left stencil wins."""
    mm, pos, x0 = window(3, di)
    f = [(x + 4) ** 3 for x in pos]
    f[2] += 1  # third differences become (2, 12, 2) for left, centred, right
    f[3] -= 1
    cands = [list(range(s, s + 4)) for s in range(3)]
    V = [undivided([f[k] for k in c]) for c in cands]
    assert V[0] == V[2] and abs(V[0]) < abs(V[1]) and V[0] * V[1] > 0
    vals = [lagrange([pos[k] for k in c], [f[k] for k in c], x0) for c in cands]
    assert len(set(vals)) == 3
    assert min(f[-mm - 1 + di], f[-mm + di]) < vals[0] < max(f[-mm - 1 + di], f[-mm + di])
    got = orc.eno1(3, di, np.array([float(v) for v in f]))
    assert F(got) == vals[0]


def kinked_axis(n, order, kinks, amp=4):
    """Values of a kinked polynomial at coarse cells 0..n-1 (centre c+1/2):
    (x+3)^ORDER plus, per kink (t, side), amp*(x-t)_+^(ORDER-1) right of t or
    its sign-matched mirror left of t."""
    x = [F(2 * c + 1, 2) for c in range(n)]
    m = order - 1
    out = []
    for xc in x:
        v = (xc + 3) ** order
        for t, side in kinks:
            if side > 0 and xc > t:
                v += amp * (xc - t) ** m
            if side < 0 and xc < t:
                v += amp * (-1) ** order * (t - xc) ** m
        out.append(v)
    return out


def _cc_box(lo_cell, n, stride):
    """bbox of n cell-centred points starting at cell lo_cell, stride s:
    centres at s/2 + c*s."""
    return ((stride // 2 + lo_cell * stride,) * 3, (stride // 2 + (lo_cell + n - 1) * stride,) * 3, (stride,) * 3)


@pytest.mark.parametrize("order", [2, 3])
def test_eno_3d_separable_kinked_sum(order):
    """A field g(x) + h(y) + k(z), each axis kinked at different places: the
    undivided differences along an axis do not see the other two terms, so
    the nested x -> y -> z operator must equal the sum of the three 1-D ENO
    values, bit for bit (all data and weights are dyadic, every sum exact).
    The 1-D values select non-centred stencils at many fine points."""
    n = 14
    gx = kinked_axis(n, order, [(F(9, 2) + 1, +1)])
    gy = kinked_axis(n, order, [(F(15, 2) + 1, -1)])
    gz = kinked_axis(n, order, [(F(11, 2) + 1, +1), (F(19, 2) + 1, -1)])
    sb = _cc_box(0, n, 4)
    src = host_array(sb, sizes(sb))
    src.view()[...] = (np.array([float(v) for v in gz])[:, None, None]
                       + np.array([float(v) for v in gy])[None, :, None]
                       + np.array([float(v) for v in gx])[None, None, :])
    # every fine cell whose stencil window lies inside the coarse box
    rc, mm0, mx0 = orc.eno_window(order, 0)
    rc, mm1, mx1 = orc.eno_window(order, 1)
    f_lo = 2 * (-min(mm0, mm1))
    f_hi = 2 * (n - max(mx0, mx1)) + 1
    nf = f_hi - f_lo + 1
    rb = _cc_box(f_lo, nf, 2)
    dst = host_array(rb, sizes(rb))
    assert orc.prolongate_3d_cc_eno_rf2(order, src, dst, rb) == 0

    def line_values(g):
        vals, moved = [], 0
        for fcell in range(f_lo, f_hi + 1):
            b, di = fcell >> 1, fcell & 1
            rc, mm, mx = orc.eno_window(order, di)
            win = [g[b + i] for i in range(mm, mx)]
            v = orc.eno1(order, di, np.array([float(w) for w in win]))
            # centred (ORDER 3) / unshifted (ORDER 2) stencil's value
            pos = [F(2 * i + 1, 2) for i in range(mm, mx)]
            x0 = F(1, 4) + F(di, 2)
            c0 = 1 if order == 3 else (1 if di == 0 else 0)
            nodes = list(range(c0, c0 + order + 1))
            plain = lagrange([pos[k] for k in nodes], [win[k] for k in nodes], x0)
            moved += F(v) != plain
            vals.append(F(v))
        return vals, moved

    ex, mx_ = line_values(gx)
    ey, my_ = line_values(gy)
    ez, mz_ = line_values(gz)
    assert mx_ >= 2 and my_ >= 2 and mz_ >= 4, (mx_, my_, mz_)
    got = dst.view()
    for k in range(nf):
        for j in range(nf):
            for i in range(nf):
                assert F(got[k, j, i]) == ex[i] + ey[j] + ez[k], (i, j, k)


@pytest.mark.parametrize("order", [2, 3])
def test_eno_3d_nests_x_inside_y_inside_z(order):
    """On random data the 3-D operator equals 1-D ENO along x on every
    (j, k) row of the window, then along y on those results, then along z
    (interp3 -> interp2 -> interp1, :943-1363); taking the axes in the
    opposite order gives different values, so the test sees the order."""
    rng = np.random.default_rng(50 + order)
    n = 10
    sb = _cc_box(0, n, 4)
    src = host_array(sb, sizes(sb))
    v = rng.uniform(-1, 1, size=(n, n, n))
    src.view()[...] = v
    rb = _cc_box(6, 6, 2)  # fine cells 6..11: bases 3..5
    dst = host_array(rb, sizes(rb))
    assert orc.prolongate_3d_cc_eno_rf2(order, src, dst, rb) == 0

    def eno(line, di):
        return orc.eno1(order, di, np.ascontiguousarray(line))

    def nested(vol, p, axes):
        # vol indexed [z, y, x]; p = fine cell per axis (x, y, z)
        rngs = []
        for d in range(3):
            b, di = p[d] >> 1, p[d] & 1
            rc, mm, mx = orc.eno_window(order, di)
            rngs.append((b, di, mm, mx))
        (bx, dx, mx0, mx1), (by, dy, my0, my1), (bz, dz, mz0, mz1) = rngs
        blk = vol[bz + mz0:bz + mz1, by + my0:by + my1, bx + mx0:bx + mx1]
        for ax in axes:  # 2 = x, 1 = y, 0 = z (numpy axis of blk)
            di = {2: dx, 1: dy, 0: dz}[ax]
            blk = np.apply_along_axis(eno, ax, blk, di)
            blk = np.expand_dims(blk, ax)
        return float(blk.reshape(()))

    differs = 0
    for k in range(6):
        for j in range(6):
            for i in range(6):
                p = (6 + i, 6 + j, 6 + k)
                xyz = nested(v, p, (2, 1, 0))
                assert dst.view()[k, j, i] == xyz, (i, j, k)
                differs += nested(v, p, (0, 1, 2)) != xyz
    assert differs > 0
