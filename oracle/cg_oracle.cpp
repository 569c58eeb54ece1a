// cg_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of CarpetLib's inter-grid transfer operators, used as the
// checker for the HIP kernels of libcarpet_amd.so and as the timed CPU
// baseline.  Nothing in the product links, loads or calls this file; only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do.
//
// Each operator follows the reference loop structure point by point (no
// reuse of partial sums across points), so it is an independent restatement
// of what the GPU computes with separable passes.  Arithmetic policy (the
// canonical order, applied identically on the GPU): every "res += c*x" is
// res = fma(c, x, res) starting from res = 0, left to right over the
// stencil; a coincident vertex-centred direction passes the value through;
// this file is compiled with -ffp-contract=off so nothing else is fused.
//
// Pins (see tests/): the reference's own coefficient self-tests, the
// CarpetProlongateTest goldens (errornorm and 1-D lines of the difference
// to the exact polynomial), closed forms (polynomial reproduction,
// injection/copy identity, conservation of restrict_cc o prolong_tvd,
// Lagrange weights in time) and brute force on tiny inputs.
//
// Reference files followed (paths relative to the Carpet tree):
//   CarpetLib/src/copy_3d.cc:25-122
//   CarpetLib/src/prolongate_3d_rf2.cc:27-124, 151-295, 356-675
//   CarpetLib/src/prolongate_3d_cc_rf2.cc:49-195, 238-579
//   CarpetLib/src/prolongate_3d_cc_eno_rf2.cc:51-881, 943-1363
//   CarpetLib/src/prolongate_3d_cc_real8_tvd.F90:55-244
//   CarpetLib/src/restrict_3d_rf2.cc:25-121
//   CarpetLib/src/restrict_3d_cc_rf2.cc:23-152
//   CarpetLib/src/restrict_3d_vc_rf2.cc:26-260
//   CarpetLib/src/prolongate_3d_stagger{011,101,110,111}.cc:24-561,
//     coeffs_prolongate_3d_stagger.hh; restrict_3d_stagger{...}.cc:26-153,
//     coeffs_restrict_3d_stagger.hh
//   CarpetLib/src/interpolate_3d_2tl.cc:24-117, interpolate_3d_3tl.cc:24-126
//   CarpetLib/src/data.cc:34-105 (call_operator), 485-534 (transfer_time)
//   CarpetLib/src/gdata.cc:288-349 (find_source_timelevel)

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "carpet_amd.h" // error-code and enum constants only

namespace {

struct box3 {
  int lo[3], up[3], str[3];
};

box3 mkbox(const int *b) {
  box3 r;
  for (int d = 0; d < 3; ++d) {
    r.lo[d] = b[d];
    r.up[d] = b[3 + d];
    r.str[d] = b[6 + d];
  }
  return r;
}

// bbox<int,3> semantics, CarpetLib/src/bbox.cc:64-126, 181-191
bool empty(const box3 &b) {
  for (int d = 0; d < 3; ++d)
    if (b.lo[d] > b.up[d])
      return true;
  return false;
}
int shape(const box3 &b, int d) { return empty(b) ? 0 : b.up[d] - b.lo[d] + b.str[d]; }
int sizes(const box3 &b, int d) { return shape(b, d) / b.str[d]; }
bool contained(const box3 &a, const box3 &b) {
  if (empty(a))
    return true;
  for (int d = 0; d < 3; ++d)
    if (a.lo[d] < b.lo[d] || a.up[d] > b.up[d])
      return false;
  return true;
}

inline size_t idx3(ptrdiff_t i, ptrdiff_t j, ptrdiff_t k, const int *pad) {
  return (size_t)(i + (ptrdiff_t)pad[0] * (j + (ptrdiff_t)pad[1] * k));
}

// Common array-shape check: ext must be the box sizes and fit the padding.
int check_array(const int *pad, const int *ext, const box3 &b) {
  for (int d = 0; d < 3; ++d) {
    if (b.str[d] <= 0)
      return CG_ERR_ARG;
    if (ext[d] != sizes(b, d) || pad[d] < ext[d])
      return CG_ERR_SHAPE;
  }
  return CG_OK;
}

// Lagrange coefficients -- CarpetLib/src/prolongate_3d_rf2.cc:92-124 (vertex)
const double vc_c1[] = {+1 / 2.0, +1 / 2.0};
const double vc_c3[] = {-1 / 16.0, +9 / 16.0, +9 / 16.0, -1 / 16.0};
const double vc_c5[] = {+3 / 256.0, -25 / 256.0, +150 / 256.0,
                        +150 / 256.0, -25 / 256.0, +3 / 256.0};
const double vc_c7[] = {-5 / 2048.0, +49 / 2048.0, -245 / 2048.0,
                        +1225 / 2048.0, +1225 / 2048.0, -245 / 2048.0,
                        +49 / 2048.0, -5 / 2048.0};
const double vc_c9[] = {+35 / 65536.0, -405 / 65536.0, +2268 / 65536.0,
                        -8820 / 65536.0, +39690 / 65536.0, +39690 / 65536.0,
                        -8820 / 65536.0, +2268 / 65536.0, -405 / 65536.0,
                        +35 / 65536.0};
const double vc_c11[] = {-63 / 524288.0, +847 / 524288.0, -5445 / 524288.0,
                         +22869 / 524288.0, -76230 / 524288.0,
                         +320166 / 524288.0, +320166 / 524288.0,
                         -76230 / 524288.0, +22869 / 524288.0,
                         -5445 / 524288.0, +847 / 524288.0, -63 / 524288.0};
const double *vc_coeffs(int order) {
  switch (order) {
  case 1: return vc_c1;
  case 3: return vc_c3;
  case 5: return vc_c5;
  case 7: return vc_c7;
  case 9: return vc_c9;
  case 11: return vc_c11;
  }
  return nullptr;
}

// Cell-centred Lagrange coefficients, CarpetLib/src/prolongate_3d_cc_rf2.cc:116-139
const double cc_c0[] = {+1 / 1.0};
const double cc_c1[] = {+1 / 4.0, +3 / 4.0};
const double cc_c2[] = {+5 / 32.0, +30 / 32.0, -3 / 32.0};
const double cc_c3[] = {-5 / 128.0, +35 / 128.0, +105 / 128.0, -7 / 128.0};
const double cc_c4[] = {-45 / 2048.0, +420 / 2048.0, +1890 / 2048.0,
                        -252 / 2048.0, +35 / 2048.0};
const double cc_c5[] = {+63 / 8192.0, -495 / 8192.0, +2310 / 8192.0,
                        +6930 / 8192.0, -693 / 8192.0, +77 / 8192.0};
const double *cc_coeffs(int order) {
  switch (order) {
  case 0: return cc_c0;
  case 1: return cc_c1;
  case 2: return cc_c2;
  case 3: return cc_c3;
  case 4: return cc_c4;
  case 5: return cc_c5;
  }
  return nullptr;
}

// ---- 1-D stencil descriptions --------------------------------------------
// A direction of a tensor-product prolongation is described, for each
// parity p of the fine point (0: "di=0", 1: "di=1"), by the first tap offset
// relative to the coarse base index, the number of taps and the weights; a
// vertex-centred p=0 is a pass-through.

struct Stencil1D {
  bool pass[2];
  int tmin[2];
  int ntap[2];
  double c[2][12];
};

// vertex: coeffs1d<RT,ORDER>, prolongate_3d_rf2.cc:27-46 (imin=-ncoeffs/2+1)
Stencil1D vc_stencil(int order) {
  Stencil1D s{};
  const double *c = vc_coeffs(order);
  int n = order + 1;
  s.pass[0] = true;
  s.tmin[0] = 0;
  s.ntap[0] = 1;
  s.c[0][0] = 1.0;
  s.pass[1] = false;
  s.tmin[1] = -n / 2 + 1;
  s.ntap[1] = n;
  for (int t = 0; t < n; ++t)
    s.c[1][t] = c[t];
  return s;
}

// cell: coeffs1d<RT,ORDER,di>, prolongate_3d_cc_rf2.cc:53-72
//   imin = -ncoeffs/2 + (ORDER odd ? di : 0); di==1 mirrors the weights
Stencil1D cc_stencil(int order) {
  Stencil1D s{};
  const double *c = cc_coeffs(order);
  int n = order + 1;
  for (int di = 0; di < 2; ++di) {
    s.pass[di] = false;
    s.tmin[di] = -n / 2 + (order % 2 != 0 ? di : 0);
    s.ntap[di] = n;
    for (int t = 0; t < n; ++t)
      s.c[di][t] = di == 0 ? c[t] : c[n - 1 - t];
  }
  return s;
}

// One 1-D Lagrange sum over values v[0..ntap-1] (already the taps), in the
// canonical order.
inline double lag1(const Stencil1D &s, int p, const double *v) {
  if (s.pass[p])
    return v[0];
  double res = 0.0;
  for (int t = 0; t < s.ntap[p]; ++t)
    res = std::fma(s.c[p][t], v[t], res);
  return res;
}

// Tensor-product nested evaluation of one fine point: interp3 -> interp2 ->
// interp1 (prolongate_3d_rf2.cc:158-295, prolongate_3d_cc_rf2.cc:157-195).
double lag3(const Stencil1D &s, const double *src, const int *pad, ptrdiff_t is,
            ptrdiff_t js, ptrdiff_t ks, int pi, int pj, int pk) {
  double vz[12];
  for (int tk = 0; tk < s.ntap[pk]; ++tk) {
    ptrdiff_t k = ks + s.tmin[pk] + tk;
    double vy[12];
    for (int tj = 0; tj < s.ntap[pj]; ++tj) {
      ptrdiff_t j = js + s.tmin[pj] + tj;
      double vx[12];
      for (int ti = 0; ti < s.ntap[pi]; ++ti)
        vx[ti] = src[idx3(is + s.tmin[pi] + ti, j, k, pad)];
      vy[tj] = lag1(s, pi, vx);
    }
    vz[tk] = lag1(s, pj, vy);
  }
  return lag1(s, pk, vz);
}

// ---- ENO (cell-centred) 1-D operator ---------------------------------------
// coeffs1d<RT,ORDER,di,OFFSET>, prolongate_3d_cc_eno_rf2.cc:57-99, 144-174
// coeffs1dc<RT,ORDER,OFFSET> with the OFFSET taken as (di==0 ? OFFSET : -OFFSET)
const double eno_c1_0[] = {+1 / 4.0, +3 / 4.0};
const double eno_c2_0[] = {+5 / 32.0, +30 / 32.0, -3 / 32.0};
const double eno_c2_m1[] = {-0.09375, +0.4375, +0.65625};
const double eno_c2_p1[] = {0, 0, 0};
const double eno_c3_0[] = {-0.0390625, 0.2734375, 0.8203125, -0.0546875};
const double eno_c3_p1[] = {0.1171875, 1.0546875, -0.2109375, 0.0390625};
const double eno_c3_m1[] = {0.0546875, -0.2578125, 0.6015625, 0.6015625};
const double eno_d2[] = {+1.0, -2.0, +1.0};
const double eno_d3[] = {+1.0, -3.0, +3.0, -1.0};

const double *eno_table(int order, int offset) {
  if (order == 1)
    return eno_c1_0;
  if (order == 2)
    return offset == 0 ? eno_c2_0 : offset < 0 ? eno_c2_m1 : eno_c2_p1;
  if (offset == 0)
    return eno_c3_0;
  return offset > 0 ? eno_c3_p1 : eno_c3_m1;
}

struct EnoCoeffs {
  int n, imin, imax;
  int minimin, maximax;
  const double *tab;
  int di;
  double get(int i) const {
    int j = di == 0 ? i - imin : n - 1 - (i - imin);
    return tab[j];
  }
};

EnoCoeffs eno_coeffs(int order, int di, int offset) {
  EnoCoeffs e;
  e.n = order + 1;
  e.imin = -e.n / 2 + (order % 2 != 0 ? di : 0) + offset;
  e.imax = e.imin + e.n;
  e.minimin = e.imin - order / 2 + (order % 2 == 0 ? di : 0) - offset;
  e.maximax = e.imax + order / 2 - (order % 2 == 0 ? 1 - di : 0) - offset;
  e.tab = eno_table(order, di == 0 ? offset : -offset);
  e.di = di;
  return e;
}

// The body shared by interp1/interp2/interp3 of prolongate_3d_cc_eno_rf2.cc
// (:207-416, :435-641, :661-881): f[] holds the values at offsets
// minimin..maximax-1 of the coarse base.
double eno1(int order, int di, const double *f) {
  const EnoCoeffs c0 = eno_coeffs(order, di, 0);
  const int mm = c0.minimin;
  const EnoCoeffs lin = eno_coeffs(1, di, 0);
  auto linear = [&]() {
    double r = 0.0;
    for (int i = lin.imin; i < lin.imax; ++i)
      r = std::fma(lin.get(i), f[i - mm], r);
    return r;
  };
  double res = 0.0;
  if (order == 2) {
    const int shift = di == 0 ? -1 : 1;
    const EnoCoeffs rc = eno_coeffs(2, di, 0);
    const EnoCoeffs lc = eno_coeffs(2, di, shift);
    double lV = 0.0, rV = 0.0;
    for (int i = lc.imin; i < lc.imax; ++i)
      lV = std::fma(eno_d2[i - lc.imin], f[i - mm], lV);
    for (int i = rc.imin; i < rc.imax; ++i)
      rV = std::fma(eno_d2[i - rc.imin], f[i - mm], rV);
    if (lV * rV <= 0) {
      res = linear(); // falls through to the monotonicity check (:279)
    } else if (std::fabs(lV) < std::fabs(rV)) {
      for (int i = lc.imin; i < lc.imax; ++i)
        res = std::fma(lc.get(i), f[i - mm], res);
    } else {
      for (int i = rc.imin; i < rc.imax; ++i)
        res = std::fma(rc.get(i), f[i - mm], res);
    }
  } else { // order 3
    const EnoCoeffs lc = eno_coeffs(3, di, -1);
    const EnoCoeffs cc = eno_coeffs(3, di, 0);
    const EnoCoeffs rc = eno_coeffs(3, di, +1);
    double V[3] = {0.0, 0.0, 0.0};
    for (int i = lc.imin; i < lc.imax; ++i)
      V[0] = std::fma(eno_d3[i - lc.imin], f[i - mm], V[0]);
    for (int i = cc.imin; i < cc.imax; ++i)
      V[1] = std::fma(eno_d3[i - cc.imin], f[i - mm], V[1]);
    for (int i = rc.imin; i < rc.imax; ++i)
      V[2] = std::fma(eno_d3[i - rc.imin], f[i - mm], V[2]);
    if (V[0] * V[2] <= 0 || V[0] * V[1] <= 0 || V[1] * V[2] <= 0)
      return linear(); // returns without the monotonicity check (:338)
    int min = 1;
    for (int i = 0; i < 3; ++i)
      if (std::fabs(V[i]) < std::fabs(V[min]))
        min = i;
    const EnoCoeffs &sel = min == 0 ? lc : min == 1 ? cc : rc;
    for (int i = sel.imin; i < sel.imax; ++i)
      res = std::fma(sel.get(i), f[i - mm], res);
  }
  // monotonicity check against the two nearest coarse values (:378-408)
  if ((res - f[-mm - 1 + di]) * (f[-mm + di] - res) < 0)
    res = linear();
  return res;
}

double eno3(int order, const double *src, const int *pad, ptrdiff_t is,
            ptrdiff_t js, ptrdiff_t ks, int pi, int pj, int pk) {
  const EnoCoeffs ci = eno_coeffs(order, pi, 0);
  const EnoCoeffs cj = eno_coeffs(order, pj, 0);
  const EnoCoeffs ck = eno_coeffs(order, pk, 0);
  double fz[8];
  for (int tk = ck.minimin; tk < ck.maximax; ++tk) {
    double fy[8];
    for (int tj = cj.minimin; tj < cj.maximax; ++tj) {
      double fx[8];
      for (int ti = ci.minimin; ti < ci.maximax; ++ti)
        fx[ti - ci.minimin] = src[idx3(is + ti, js + tj, ks + tk, pad)];
      fy[tj - cj.minimin] = eno1(order, pi, fx);
    }
    fz[tk - ck.minimin] = eno1(order, pj, fy);
  }
  return eno1(order, pk, fz);
}

// ---- call_operator: OpenMP split along the longest axis (data.cc:56-100) --
template <typename F>
void split_longest(const box3 &reg, int nthreads_hint, F &&body) {
  int dir = 0;
  for (int d = 1; d < 3; ++d)
    if (shape(reg, d) > shape(reg, dir))
      dir = d;
  const int stride = reg.str[dir];
  const int first = reg.lo[dir];
  const int last = reg.up[dir] + stride;
  const int npts = last - first;
#pragma omp parallel
  {
    int nth = 1, me = 0;
#ifdef _OPENMP
    nth = omp_get_num_threads();
    me = omp_get_thread_num();
#endif
    (void)nthreads_hint;
    int mynum = (npts / stride + nth - 1) / nth * stride;
    int myfirst = std::min(last, first + me * mynum);
    int mylast = std::min(last, myfirst + mynum);
    box3 sub = reg;
    sub.lo[dir] = myfirst;
    sub.up[dir] = mylast - stride;
    if (!empty(sub))
      body(sub);
  }
}

int g_use_openmp = 1;

// eno1d of prolongate_3d_real8_eno.F90:8-67, each operation rounded on its
// own in the order written
double eno1d(const double *q) {
  const double diffleft = (q[0] + q[2]) - 2.0 * q[1];
  const double diffright = (q[1] + q[3]) - 2.0 * q[2];
  double r;
  if (std::fabs(diffleft) < std::fabs(diffright))
    r = 0.125 * ((-q[0] + 6.0 * q[1]) + 3.0 * q[2]);
  else
    r = 0.125 * ((3.0 * q[1] + 6.0 * q[2]) - q[3]);
  if ((r - q[1]) * (q[2] - r) < 0.0 || diffleft * diffright <= 0.0)
    r = 0.5 * (q[1] + q[2]);
  return r;
}

// ---- operators -----------------------------------------------------------

int prolong_tensor(bool cell, int order, const double *src, const int *spad,
                   const int *sext, double *dst, const int *dpad,
                   const int *dext, const int *sb, const int *db,
                   const int *rb) {
  const box3 srcbbox = mkbox(sb), dstbbox = mkbox(db), reg = mkbox(rb);
  int e;
  if ((e = check_array(spad, sext, srcbbox)) || (e = check_array(dpad, dext, dstbbox)))
    return e;
  if (reg.str[0] <= 0 || reg.str[1] <= 0 || reg.str[2] <= 0)
    return CG_ERR_ARG;
  for (int d = 0; d < 3; ++d) {
    if (srcbbox.str[d] <= reg.str[d] || dstbbox.str[d] != reg.str[d])
      return CG_ERR_STRIDE;
    if (srcbbox.str[d] != 2 * dstbbox.str[d])
      return CG_ERR_REFFACT;
  }
  if (!cell) {
    for (int d = 0; d < 3; ++d)
      if (srcbbox.lo[d] % srcbbox.str[d] || dstbbox.lo[d] % dstbbox.str[d] ||
          reg.lo[d] % reg.str[d])
        return CG_ERR_ALIGN;
  } else {
    for (int d = 0; d < 3; ++d)
      if (dstbbox.str[d] % 2)
        return CG_ERR_ALIGN;
  }
  if (empty(reg))
    return CG_ERR_EMPTY;
  if (!contained(reg, dstbbox))
    return CG_ERR_CONTAIN;
  Stencil1D st = cell ? cc_stencil(order) : vc_stencil(order);
  int off[3], doff[3], n[3];
  for (int d = 0; d < 3; ++d) {
    int num = reg.lo[d] - srcbbox.lo[d] + (cell ? reg.str[d] / 2 : 0);
    if (num % reg.str[d] || (reg.lo[d] - dstbbox.lo[d]) % reg.str[d])
      return CG_ERR_ALIGN;
    off[d] = num / reg.str[d];
    doff[d] = (reg.lo[d] - dstbbox.lo[d]) / reg.str[d];
    n[d] = sizes(reg, d);
    // exact stencil footprint must lie in the source array
    int lo = 1 << 30, hi = -(1 << 30);
    for (int f : {0, 1, n[d] - 2, n[d] - 1}) {
      if (f < 0 || f >= n[d])
        continue;
      int g = off[d] + f, b = g >> 1, p = g & 1;
      lo = std::min(lo, b + st.tmin[p]);
      hi = std::max(hi, b + st.tmin[p] + st.ntap[p] - 1);
    }
    if (off[d] < 0 || lo < 0 || hi >= sext[d])
      return CG_ERR_CONTAIN;
  }
  auto body = [&](const box3 &sub) {
    int s0[3], m[3];
    for (int d = 0; d < 3; ++d) {
      s0[d] = (sub.lo[d] - reg.lo[d]) / reg.str[d];
      m[d] = sizes(sub, d);
    }
    for (int k = s0[2]; k < s0[2] + m[2]; ++k)
      for (int j = s0[1]; j < s0[1] + m[1]; ++j)
        for (int i = s0[0]; i < s0[0] + m[0]; ++i) {
          int gi = off[0] + i, gj = off[1] + j, gk = off[2] + k;
          dst[idx3(doff[0] + i, doff[1] + j, doff[2] + k, dpad)] =
              lag3(st, src, spad, gi >> 1, gj >> 1, gk >> 1, gi & 1, gj & 1, gk & 1);
        }
  };
  if (g_use_openmp)
    split_longest(reg, 0, body);
  else
    body(reg);
  return CG_OK;
}

} // namespace

extern "C" {

void orc_set_openmp(int on) { g_use_openmp = on; }

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// coeffs1d<RT,ORDER>::test(), prolongate_3d_rf2.cc:49-89: sum_i c_i x_i^n
// must reproduce x0^n for n <= ORDER (x_i = i - 1/2, x0 = 0) within 1e3 eps.
// Returns the largest |res - y0| over n, or a negative value if the order
// is unknown.
double orc_vc_coeff_selftest(int order) {
  const double *c = vc_coeffs(order);
  if (!c)
    return -1;
  int n = order + 1, imin = -n / 2 + 1;
  double worst = 0;
  for (int p = 0; p <= order; ++p) {
    double res = 0;
    for (int i = imin; i < imin + n; ++i)
      res += c[i - imin] * std::pow(double(i) - 0.5, p);
    worst = std::max(worst, std::fabs(res - (p == 0 ? 1.0 : 0.0)));
  }
  return worst;
}

// coeffs1d<RT,ORDER,di>::test(), prolongate_3d_cc_rf2.cc:75-113:
// x_i = i + 1/2, x0 = 1/4 + di/2.
double orc_cc_coeff_selftest(int order, int di) {
  Stencil1D s = cc_stencil(order);
  if (!cc_coeffs(order) || di < 0 || di > 1)
    return -1;
  double worst = 0;
  for (int p = 0; p <= order; ++p) {
    double res = 0;
    for (int t = 0; t < s.ntap[di]; ++t)
      res += s.c[di][t] * std::pow(double(s.tmin[di] + t) + 0.5, p);
    worst = std::max(worst, std::fabs(res - std::pow(0.25 + 0.5 * di, p)));
  }
  return worst;
}

// coeffs1d<RT,ORDER,di,OFFSET>::test(), prolongate_3d_cc_eno_rf2.cc:102-141
double orc_eno_coeff_selftest(int order, int di, int offset) {
  if (order < 1 || order > 3 || di < 0 || di > 1)
    return -1;
  EnoCoeffs e = eno_coeffs(order, di, offset);
  double worst = 0;
  for (int p = 0; p <= order; ++p) {
    double res = 0;
    for (int i = e.imin; i < e.imax; ++i)
      res += e.get(i) * std::pow(double(i) + 0.5, p);
    worst = std::max(worst, std::fabs(res - std::pow(0.25 + 0.5 * di, p)));
  }
  return worst;
}

// ENO 1-D operator on an explicit window f[minimin..maximax-1] (exported so
// tests can check its branch logic on hand-made lines)
double orc_eno1(int order, int di, const double *f) { return eno1(order, di, f); }

int orc_eno_window(int order, int di, int *minimin, int *maximax) {
  if (order < 2 || order > 3 || di < 0 || di > 1)
    return CG_ERR_ORDER;
  EnoCoeffs e = eno_coeffs(order, di, 0);
  *minimin = e.minimin;
  *maximax = e.maximax;
  return CG_OK;
}

// copy_3d, copy_3d.cc:25-122
int orc_copy_3d(const double *src, const int *spad, const int *sext,
                double *dst, const int *dpad, const int *dext, const int *sb,
                const int *db, const int *srb, const int *drb,
                const int *slab_offset) {
  const box3 srcbbox = mkbox(sb), dstbbox = mkbox(db), srcreg1 = mkbox(srb),
             dreg = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, srcbbox)) || (e = check_array(dpad, dext, dstbbox)))
    return e;
  for (int d = 0; d < 3; ++d)
    if (srcbbox.str[d] != dreg.str[d] || dstbbox.str[d] != dreg.str[d] ||
        srcreg1.str[d] != dreg.str[d])
      return CG_ERR_STRIDE;
  if (empty(dreg))
    return CG_ERR_EMPTY;
  box3 sreg = dreg;
  if (slab_offset)
    for (int d = 0; d < 3; ++d) {
      sreg.lo[d] += slab_offset[d] * dreg.str[d];
      sreg.up[d] += slab_offset[d] * dreg.str[d];
    }
  if (!contained(sreg, srcreg1))
    return CG_ERR_CONTAIN;
  if (!contained(sreg, srcbbox) || !contained(dreg, dstbbox))
    return CG_ERR_CONTAIN;
  int so[3], dof[3], n[3];
  for (int d = 0; d < 3; ++d) {
    if ((sreg.lo[d] - srcbbox.lo[d]) % srcbbox.str[d] ||
        (dreg.lo[d] - dstbbox.lo[d]) % dstbbox.str[d])
      return CG_ERR_ALIGN;
    so[d] = (sreg.lo[d] - srcbbox.lo[d]) / srcbbox.str[d];
    dof[d] = (dreg.lo[d] - dstbbox.lo[d]) / dstbbox.str[d];
    n[d] = sizes(dreg, d);
  }
#pragma omp parallel for collapse(2) if (g_use_openmp)
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i)
        dst[idx3(dof[0] + i, dof[1] + j, dof[2] + k, dpad)] =
            src[idx3(so[0] + i, so[1] + j, so[2] + k, spad)];
  return CG_OK;
}

// prolongate_3d_rf2<double,ORDER>
int orc_prolongate_3d_rf2(int order, const double *src, const int *spad,
                          const int *sext, double *dst, const int *dpad,
                          const int *dext, const int *sb, const int *db,
                          const int *srb, const int *drb) {
  (void)srb;
  if (!vc_coeffs(order))
    return CG_ERR_ORDER;
  return prolong_tensor(false, order, src, spad, sext, dst, dpad, dext, sb, db, drb);
}

// prolongate_3d_real8_eno (vertex-centred ENO, prolongate_3d_real8_eno.F90:
// 69-288): per fine point the case of its parities (fi + 10 fj + 100 fk),
// with the F90's temporaries -- x lines first, then y, then z; a
// coincident direction takes the base value
int orc_prolongate_3d_eno(const double *src, const int *spad, const int *sext, double *dst,
                          const int *dpad, const int *dext, const int *sb, const int *db,
                          const int *srb, const int *drb) {
  (void)srb;
  const box3 srcbbox = mkbox(sb), dstbbox = mkbox(db), reg = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, srcbbox)) || (e = check_array(dpad, dext, dstbbox)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (reg.str[d] <= 0)
      return CG_ERR_ARG;
    if (srcbbox.str[d] <= reg.str[d] || dstbbox.str[d] != reg.str[d])
      return CG_ERR_STRIDE;
    if (srcbbox.str[d] != 2 * dstbbox.str[d])
      return CG_ERR_REFFACT;
    if (srcbbox.lo[d] % srcbbox.str[d] || dstbbox.lo[d] % dstbbox.str[d] || reg.lo[d] % reg.str[d])
      return CG_ERR_ALIGN;
  }
  if (empty(reg))
    return CG_ERR_EMPTY;
  if (!contained(reg, dstbbox))
    return CG_ERR_CONTAIN;
  int off[3], doff[3], n[3];
  for (int d = 0; d < 3; ++d) {
    const int num = reg.lo[d] - srcbbox.lo[d];
    if (num % reg.str[d] || (reg.lo[d] - dstbbox.lo[d]) % reg.str[d])
      return CG_ERR_ALIGN;
    off[d] = num / reg.str[d];
    doff[d] = (reg.lo[d] - dstbbox.lo[d]) / reg.str[d];
    n[d] = sizes(reg, d);
    // exact footprint: b for a coincident child, b-1..b+2 for a midpoint
    int lo = 1 << 30, hi = -(1 << 30);
    for (int f : {0, 1, n[d] - 2, n[d] - 1}) {
      if (f < 0 || f >= n[d])
        continue;
      const int g = off[d] + f, b = g >> 1;
      lo = std::min(lo, (g & 1) ? b - 1 : b);
      hi = std::max(hi, (g & 1) ? b + 2 : b);
    }
    if (off[d] < 0 || lo < 0 || hi >= sext[d])
      return CG_ERR_CONTAIN;
  }
  auto at = [&](int i, int j, int k) { return src[idx3(i, j, k, spad)]; };
  auto body = [&](const box3 &sub) {
    int s0[3], m[3];
    for (int d = 0; d < 3; ++d) {
      s0[d] = (sub.lo[d] - reg.lo[d]) / reg.str[d];
      m[d] = sizes(sub, d);
    }
    for (int k = s0[2]; k < s0[2] + m[2]; ++k)
      for (int j = s0[1]; j < s0[1] + m[1]; ++j)
        for (int i = s0[0]; i < s0[0] + m[0]; ++i) {
          const int gi = off[0] + i, gj = off[1] + j, gk = off[2] + k;
          const int i0 = gi >> 1, j0 = gj >> 1, k0 = gk >> 1;
          const int fi = gi & 1, fj = gj & 1, fk = gk & 1;
          double q[4], tmp2[4], tmp1[4][4], out = 0.0;
          switch (fi + 10 * fj + 100 * fk) {
          case 0:
            out = at(i0, j0, k0);
            break;
          case 1:
            for (int t = 0; t < 4; ++t)
              q[t] = at(i0 - 1 + t, j0, k0);
            out = eno1d(q);
            break;
          case 10:
            for (int t = 0; t < 4; ++t)
              q[t] = at(i0, j0 - 1 + t, k0);
            out = eno1d(q);
            break;
          case 11:
            for (int jj = 0; jj < 4; ++jj) {
              for (int t = 0; t < 4; ++t)
                q[t] = at(i0 - 1 + t, j0 - 1 + jj, k0);
              tmp2[jj] = eno1d(q);
            }
            out = eno1d(tmp2);
            break;
          case 100:
            for (int t = 0; t < 4; ++t)
              q[t] = at(i0, j0, k0 - 1 + t);
            out = eno1d(q);
            break;
          case 101:
            for (int kk = 0; kk < 4; ++kk) {
              for (int t = 0; t < 4; ++t)
                q[t] = at(i0 - 1 + t, j0, k0 - 1 + kk);
              tmp2[kk] = eno1d(q);
            }
            out = eno1d(tmp2);
            break;
          case 110:
            for (int kk = 0; kk < 4; ++kk) {
              for (int t = 0; t < 4; ++t)
                q[t] = at(i0, j0 - 1 + t, k0 - 1 + kk);
              tmp2[kk] = eno1d(q);
            }
            out = eno1d(tmp2);
            break;
          default: // 111
            for (int jj = 0; jj < 4; ++jj)
              for (int kk = 0; kk < 4; ++kk) {
                for (int t = 0; t < 4; ++t)
                  q[t] = at(i0 - 1 + t, j0 - 1 + jj, k0 - 1 + kk);
                tmp1[jj][kk] = eno1d(q);
              }
            for (int ii = 0; ii < 4; ++ii) {
              for (int t = 0; t < 4; ++t)
                q[t] = tmp1[t][ii];
              tmp2[ii] = eno1d(q);
            }
            out = eno1d(tmp2);
            break;
          }
          dst[idx3(doff[0] + i, doff[1] + j, doff[2] + k, dpad)] = out;
        }
  };
  if (g_use_openmp)
    split_longest(reg, 0, body);
  else
    body(reg);
  return CG_OK;
}

double orc_eno1d(const double *q) { return eno1d(q); }

// prolongate_3d_o5_monotone_rf2 (prolongate_3d_o5_monotone_rf2.cc:45-791):
// the reference's eight parity cases, each with its own term order
int orc_prolongate_3d_o5_monotone_rf2(const double *src, const int *spad, const int *sext, double *dst,
                                      const int *dpad, const int *dext, const int *sb, const int *db,
                                      const int *srb, const int *drb) {
  (void)srb;
  const box3 srcbbox = mkbox(sb), dstbbox = mkbox(db), reg = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, srcbbox)) || (e = check_array(dpad, dext, dstbbox)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (reg.str[d] <= 0)
      return CG_ERR_ARG;
    if (srcbbox.str[d] <= reg.str[d] || dstbbox.str[d] != reg.str[d])
      return CG_ERR_STRIDE;
    if (srcbbox.str[d] != 2 * dstbbox.str[d])
      return CG_ERR_REFFACT;
    if (srcbbox.lo[d] % srcbbox.str[d] || dstbbox.lo[d] % dstbbox.str[d] || reg.lo[d] % reg.str[d])
      return CG_ERR_ALIGN;
  }
  if (empty(reg))
    return CG_ERR_EMPTY;
  if (!contained(reg, dstbbox))
    return CG_ERR_CONTAIN;
  int off[3], doff[3], n[3];
  for (int d = 0; d < 3; ++d) {
    const int num = reg.lo[d] - srcbbox.lo[d];
    if (num % reg.str[d] || (reg.lo[d] - dstbbox.lo[d]) % reg.str[d])
      return CG_ERR_ALIGN;
    off[d] = num / reg.str[d];
    doff[d] = (reg.lo[d] - dstbbox.lo[d]) / reg.str[d];
    n[d] = sizes(reg, d);
    int lo = 1 << 30, hi = -(1 << 30);
    for (int f : {0, 1, n[d] - 2, n[d] - 1}) {
      if (f < 0 || f >= n[d])
        continue;
      const int g = off[d] + f, b = g >> 1;
      lo = std::min(lo, (g & 1) ? b - 2 : b);
      hi = std::max(hi, (g & 1) ? b + 3 : b);
    }
    if (off[d] < 0 || lo < 0 || hi >= sext[d])
      return CG_ERR_CONTAIN;
  }
  const double f[6] = {3.0 / 256, -25.0 / 256, 150.0 / 256, 150.0 / 256, -25.0 / 256, 3.0 / 256};
  const double h = 1.0 / 2;
  auto S = [&](int i, int j, int k) { return src[idx3(i, j, k, spad)]; };
  // one odd axis: six taps along it (cases l8001, l8010, l8100)
  auto one = [&](int is, int js, int ks, int di, int dj, int dk) {
    double r = f[0] * S(is - 2 * di, js - 2 * dj, ks - 2 * dk);
    for (int t = 1; t < 6; ++t)
      r = r + f[t] * S(is + (t - 2) * di, js + (t - 2) * dj, ks + (t - 2) * dk);
    const double a = S(is, js, ks), b = S(is + di, js + dj, ks + dk);
    if (r > std::max(a, b) || r < std::min(a, b))
      r = h * a + h * b;
    return r;
  };
  // two odd axes, u inner and v outer (cases l8011: x,y; l8101: x,z;
  // l8110: y,z)
  auto two = [&](int is, int js, int ks, const int *u, const int *v) {
    double r = 0.0;
    bool first = true;
    for (int tv = 0; tv < 6; ++tv)
      for (int tu = 0; tu < 6; ++tu) {
        const double w = f[tu] * f[tv];
        const double x = S(is + (tu - 2) * u[0] + (tv - 2) * v[0], js + (tu - 2) * u[1] + (tv - 2) * v[1],
                           ks + (tu - 2) * u[2] + (tv - 2) * v[2]);
        r = first ? w * x : r + w * x;
        first = false;
      }
    const double c00 = S(is, js, ks), c10 = S(is + u[0], js + u[1], ks + u[2]),
                 c01 = S(is + v[0], js + v[1], ks + v[2]),
                 c11 = S(is + u[0] + v[0], js + u[1] + v[1], ks + u[2] + v[2]);
    const double mx = std::max(std::max(c00, c10), std::max(c01, c11));
    const double mn = std::min(std::min(c00, c10), std::min(c01, c11));
    if (r > mx || r < mn)
      r = h * h * c00 + h * h * c10 + h * h * c01 + h * h * c11;
    return r;
  };
  // three odd axes (case l8111): res1..res6 per z offset, then their sum
  auto three = [&](int is, int js, int ks) {
    double res[6];
    for (int c = 0; c < 6; ++c) {
      double r = 0.0;
      bool first = true;
      for (int b = 0; b < 6; ++b)
        for (int a = 0; a < 6; ++a) {
          const double w = f[a] * f[b] * f[c];
          const double x = S(is + a - 2, js + b - 2, ks + c - 2);
          r = first ? w * x : r + w * x;
          first = false;
        }
      res[c] = r;
    }
    double r = res[0] + res[1] + res[2] + res[3] + res[4] + res[5];
    const double c[8] = {S(is, js, ks),         S(is + 1, js, ks),     S(is, js + 1, ks),
                         S(is, js, ks + 1),     S(is + 1, js + 1, ks), S(is + 1, js, ks + 1),
                         S(is, js + 1, ks + 1), S(is + 1, js + 1, ks + 1)};
    const double mx = std::max(std::max(std::max(c[0], c[1]), std::max(c[2], c[3])),
                               std::max(std::max(c[4], c[5]), std::max(c[6], c[7])));
    const double mn = std::min(std::min(std::min(c[0], c[1]), std::min(c[2], c[3])),
                               std::min(std::min(c[4], c[5]), std::min(c[6], c[7])));
    if (r > mx || r < mn) {
      const double w = h * h * h;
      const double r1 = w * S(is, js, ks) + w * S(is + 1, js, ks) + w * S(is, js + 1, ks) +
                        w * S(is + 1, js + 1, ks);
      const double r2 = w * S(is, js, ks + 1) + w * S(is + 1, js, ks + 1) + w * S(is, js + 1, ks + 1) +
                        w * S(is + 1, js + 1, ks + 1);
      r = r1 + r2;
    }
    return r;
  };
  const int X[3] = {1, 0, 0}, Y[3] = {0, 1, 0}, Z[3] = {0, 0, 1};
  auto body = [&](const box3 &sub) {
    int s0[3], m[3];
    for (int d = 0; d < 3; ++d) {
      s0[d] = (sub.lo[d] - reg.lo[d]) / reg.str[d];
      m[d] = sizes(sub, d);
    }
    for (int k = s0[2]; k < s0[2] + m[2]; ++k)
      for (int j = s0[1]; j < s0[1] + m[1]; ++j)
        for (int i = s0[0]; i < s0[0] + m[0]; ++i) {
          const int gi = off[0] + i, gj = off[1] + j, gk = off[2] + k;
          const int is = gi >> 1, js = gj >> 1, ks = gk >> 1;
          double out;
          switch ((gi & 1) + 10 * (gj & 1) + 100 * (gk & 1)) {
          case 0: out = S(is, js, ks); break;
          case 1: out = one(is, js, ks, 1, 0, 0); break;
          case 10: out = one(is, js, ks, 0, 1, 0); break;
          case 100: out = one(is, js, ks, 0, 0, 1); break;
          case 11: out = two(is, js, ks, X, Y); break;
          case 101: out = two(is, js, ks, X, Z); break;
          case 110: out = two(is, js, ks, Y, Z); break;
          default: out = three(is, js, ks); break;
          }
          dst[idx3(doff[0] + i, doff[1] + j, doff[2] + k, dpad)] = out;
        }
  };
  if (g_use_openmp)
    split_longest(reg, 0, body);
  else
    body(reg);
  return CG_OK;
}

// prolongate_3d_cc_rf2<double,ORDER>
int orc_prolongate_3d_cc_rf2(int order, const double *src, const int *spad,
                             const int *sext, double *dst, const int *dpad,
                             const int *dext, const int *sb, const int *db,
                             const int *srb, const int *drb) {
  (void)srb;
  if (!cc_coeffs(order))
    return CG_ERR_ORDER;
  return prolong_tensor(true, order, src, spad, sext, dst, dpad, dext, sb, db, drb);
}

// prolongate_3d_cc_eno_rf2<double,ORDER>, ORDER in {2,3}
int orc_prolongate_3d_cc_eno_rf2(int order, const double *src, const int *spad,
                                 const int *sext, double *dst, const int *dpad,
                                 const int *dext, const int *sb, const int *db,
                                 const int *srb, const int *drb) {
  (void)srb;
  if (order != 2 && order != 3)
    return CG_ERR_ORDER;
  const box3 srcbbox = mkbox(sb), dstbbox = mkbox(db), reg = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, srcbbox)) || (e = check_array(dpad, dext, dstbbox)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (reg.str[d] <= 0)
      return CG_ERR_ARG;
    if (srcbbox.str[d] <= reg.str[d] || dstbbox.str[d] != reg.str[d])
      return CG_ERR_STRIDE;
    if (srcbbox.str[d] != 2 * dstbbox.str[d])
      return CG_ERR_REFFACT;
    if (dstbbox.str[d] % 2)
      return CG_ERR_ALIGN;
  }
  if (empty(reg))
    return CG_ERR_EMPTY;
  if (!contained(reg, dstbbox))
    return CG_ERR_CONTAIN;
  int off[3], doff[3], n[3];
  for (int d = 0; d < 3; ++d) {
    int num = reg.lo[d] - srcbbox.lo[d] + reg.str[d] / 2;
    if (num % reg.str[d] || (reg.lo[d] - dstbbox.lo[d]) % reg.str[d])
      return CG_ERR_ALIGN;
    off[d] = num / reg.str[d];
    doff[d] = (reg.lo[d] - dstbbox.lo[d]) / reg.str[d];
    n[d] = sizes(reg, d);
    int lo = 1 << 30, hi = -(1 << 30);
    for (int f : {0, 1, n[d] - 2, n[d] - 1}) {
      if (f < 0 || f >= n[d])
        continue;
      int g = off[d] + f, b = g >> 1, p = g & 1;
      EnoCoeffs c = eno_coeffs(order, p, 0);
      lo = std::min(lo, b + c.minimin);
      hi = std::max(hi, b + c.maximax - 1);
    }
    if (off[d] < 0 || lo < 0 || hi >= sext[d])
      return CG_ERR_CONTAIN;
  }
  auto body = [&](const box3 &sub) {
    int s0[3], m[3];
    for (int d = 0; d < 3; ++d) {
      s0[d] = (sub.lo[d] - reg.lo[d]) / reg.str[d];
      m[d] = sizes(sub, d);
    }
    for (int k = s0[2]; k < s0[2] + m[2]; ++k)
      for (int j = s0[1]; j < s0[1] + m[1]; ++j)
        for (int i = s0[0]; i < s0[0] + m[0]; ++i) {
          int gi = off[0] + i, gj = off[1] + j, gk = off[2] + k;
          dst[idx3(doff[0] + i, doff[1] + j, doff[2] + k, dpad)] =
              eno3(order, src, spad, gi >> 1, gj >> 1, gk >> 1, gi & 1, gj & 1, gk & 1);
        }
  };
  if (g_use_openmp)
    split_longest(reg, 0, body);
  else
    body(reg);
  return CG_OK;
}

// prolongate_3d_cc_real8_tvd, prolongate_3d_cc_real8_tvd.F90:55-244.
// Fortran indices are 1-based there; the parent cell i0+1 (1-based) is the
// 0-based cell i0 here, its neighbours i0-1 and i0+1.
int orc_prolongate_3d_cc_tvd(const double *src, const int *spad, const int *sext,
                             double *dst, const int *dpad, const int *dext,
                             const int *sb, const int *db, const int *srb,
                             const int *drb) {
  (void)srb;
  const box3 s = mkbox(sb), dbx = mkbox(db), r = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, s)) || (e = check_array(dpad, dext, dbx)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (s.str[d] == 0 || dbx.str[d] == 0 || r.str[d] == 0)
      return CG_ERR_ARG;
    if (s.str[d] <= r.str[d] || dbx.str[d] != r.str[d])
      return CG_ERR_STRIDE;
    if (s.str[d] % dbx.str[d])
      return CG_ERR_REFFACT;
    if (r.lo[d] > r.up[d])
      return CG_ERR_EMPTY;
  }
  for (int d = 0; d < 3; ++d) {
    if (s.str[d] / dbx.str[d] != 2)
      return CG_ERR_REFFACT;
    if (r.str[d] % 2)
      return CG_ERR_ALIGN;
  }
  int n[3], off[3], doff[3];
  for (int d = 0; d < 3; ++d) {
    n[d] = (r.up[d] - r.lo[d]) / r.str[d] + 1;
    // note: every axis adds regbbox(1,3)/2, the x stride (F90:165-167)
    off[d] = (r.lo[d] - s.lo[d] + r.str[0] / 2) / dbx.str[d];
    doff[d] = (r.lo[d] - dbx.lo[d]) / dbx.str[d];
    if (off[d] < 0)
      return CG_ERR_CONTAIN;
    if (doff[d] < 0 || doff[d] + n[d] > dext[d])
      return CG_ERR_CONTAIN;
    // parent cells (off+f)/2 for f in region, with their two neighbours
    int plo = off[d] / 2, phi = (off[d] + n[d] - 1) / 2;
    if (plo - 1 < 0 || phi + 1 >= sext[d])
      return CG_ERR_CONTAIN;
  }
#pragma omp parallel for collapse(2) if (g_use_openmp)
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j) {
      int k0 = (off[2] + k) / 2, fk = (off[2] + k) % 2;
      int j0 = (off[1] + j) / 2, fj = (off[1] + j) % 2;
      for (int i = 0; i < n[0]; ++i) {
        int i0 = (off[0] + i) / 2, fi = (off[0] + i) % 2;
        auto S = [&](int a, int b, int c) { return src[idx3(a, b, c, spad)]; };
        auto minmod = [](double a, double b) {
          if (a * b < 0)
            return 0.0;
          return std::fabs(a) < std::fabs(b) ? a : b;
        };
        const double fourth = 1.0 / 4;
        double c = S(i0, j0, k0);
        double dlo = c - S(i0 - 1, j0, k0), dup = S(i0 + 1, j0, k0) - c;
        double sx = double(2 * fi - 1) * fourth * minmod(dlo, dup);
        dlo = c - S(i0, j0 - 1, k0);
        dup = S(i0, j0 + 1, k0) - c;
        double sy = double(2 * fj - 1) * fourth * minmod(dlo, dup);
        dlo = c - S(i0, j0, k0 - 1);
        dup = S(i0, j0, k0 + 1) - c;
        double sz = double(2 * fk - 1) * fourth * minmod(dlo, dup);
        dst[idx3(doff[0] + i, doff[1] + j, doff[2] + k, dpad)] = c + sx + sy + sz;
      }
    }
  return CG_OK;
}

// restrict_3d_rf2, restrict_3d_rf2.cc:25-121
int orc_restrict_3d_rf2(const double *src, const int *spad, const int *sext,
                        double *dst, const int *dpad, const int *dext,
                        const int *sb, const int *db, const int *srb,
                        const int *drb) {
  (void)srb;
  const box3 s = mkbox(sb), dbx = mkbox(db), r = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, s)) || (e = check_array(dpad, dext, dbx)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (r.str[d] <= 0)
      return CG_ERR_ARG;
    if (s.str[d] >= r.str[d] || dbx.str[d] != r.str[d])
      return CG_ERR_STRIDE;
    if (2 * s.str[d] != dbx.str[d])
      return CG_ERR_REFFACT;
  }
  if (empty(r))
    return CG_ERR_EMPTY;
  if (!contained(r, s) || !contained(r, dbx))
    return CG_ERR_CONTAIN;
  int n[3], so[3], dof[3];
  for (int d = 0; d < 3; ++d) {
    if ((r.lo[d] - s.lo[d]) % s.str[d] || (r.lo[d] - dbx.lo[d]) % dbx.str[d])
      return CG_ERR_ALIGN;
    n[d] = shape(r, d) / r.str[d];
    so[d] = (r.lo[d] - s.lo[d]) / s.str[d];
    dof[d] = (r.lo[d] - dbx.lo[d]) / dbx.str[d];
  }
#pragma omp parallel for collapse(2) if (g_use_openmp)
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i)
        dst[idx3(dof[0] + i, dof[1] + j, dof[2] + k, dpad)] =
            src[idx3(so[0] + 2 * i, so[1] + 2 * j, so[2] + 2 * k, spad)];
  return CG_OK;
}

// restrict_3d_cc_rf2, restrict_3d_cc_rf2.cc:23-152
int orc_restrict_3d_cc_rf2(const double *src, const int *spad, const int *sext,
                           double *dst, const int *dpad, const int *dext,
                           const int *sb, const int *db, const int *srb,
                           const int *drb) {
  (void)srb;
  const box3 s = mkbox(sb), dbx = mkbox(db), r = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, s)) || (e = check_array(dpad, dext, dbx)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (r.str[d] <= 0)
      return CG_ERR_ARG;
    if (s.str[d] >= r.str[d] || dbx.str[d] != r.str[d])
      return CG_ERR_STRIDE;
    if (2 * s.str[d] != dbx.str[d])
      return CG_ERR_REFFACT;
    if (s.str[d] % 2)
      return CG_ERR_ALIGN;
  }
  if (empty(r))
    return CG_ERR_EMPTY;
  int n[3], so[3], dof[3];
  for (int d = 0; d < 3; ++d) {
    if ((r.lo[d] - s.lo[d] - s.str[d] / 2) % s.str[d] ||
        (r.lo[d] - dbx.lo[d]) % dbx.str[d])
      return CG_ERR_ALIGN;
    n[d] = shape(r, d) / r.str[d];
    so[d] = (r.lo[d] - s.lo[d] - s.str[d] / 2) / s.str[d];
    dof[d] = (r.lo[d] - dbx.lo[d]) / dbx.str[d];
    // the 2n fine children must lie in the source array
    if (so[d] < 0 || so[d] + 2 * n[d] > sext[d])
      return CG_ERR_CONTAIN;
  }
  if (!contained(r, dbx))
    return CG_ERR_CONTAIN;
  const double f1 = 1.0 / 2, f2 = 1.0 / 2;
#pragma omp parallel for collapse(2) if (g_use_openmp)
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        auto S = [&](int a, int b, int c) {
          return src[idx3(so[0] + a, so[1] + b, so[2] + c, spad)];
        };
        int I = 2 * i, J = 2 * j, K = 2 * k;
        double res = f1 * f1 * f1 * S(I, J, K);
        res = res + f2 * f1 * f1 * S(I + 1, J, K);
        res = res + f1 * f2 * f1 * S(I, J + 1, K);
        res = res + f2 * f2 * f1 * S(I + 1, J + 1, K);
        res = res + f1 * f1 * f2 * S(I, J, K + 1);
        res = res + f2 * f1 * f2 * S(I + 1, J, K + 1);
        res = res + f1 * f2 * f2 * S(I, J + 1, K + 1);
        res = res + f2 * f2 * f2 * S(I + 1, J + 1, K + 1);
        dst[idx3(dof[0] + i, dof[1] + j, dof[2] + k, dpad)] = res;
      }
  return CG_OK;
}

// restrict_3d_vc_rf2<T, centi, centj, centk>, restrict_3d_vc_rf2.cc:26-260:
// restriction of a cell-centred group's slab whose is_centered vector has
// one vertex-centred direction (data.cc:1106-1127; cmask bit d set =
// direction d cell-centred).  Per coarse point the nested half-sums of
// restrict1/2/3 (:26-97): a cell-centred direction averages its two
// children p and p + d (half*a + half*b), a vertex-centred one takes p;
// x innermost, then y, then z.  The children start at fine index
// srcoff + 2i, srcoff = (reg.lo - src.lo - src.str/2) / src.str (:161-171).
static double vc_restrict1(const double *p, long long d1, int cx) {
  const double half = 1.0 / 2;
  return cx ? half * p[0] + half * p[d1] : p[0];
}
static double vc_restrict2(const double *p, long long d1, long long d2, int cx, int cy) {
  const double half = 1.0 / 2;
  return cy ? half * vc_restrict1(p, d1, cx) + half * vc_restrict1(p + d2, d1, cx) : vc_restrict1(p, d1, cx);
}
static double vc_restrict3(const double *p, long long d1, long long d2, long long d3, int cx, int cy, int cz) {
  const double half = 1.0 / 2;
  return cz ? half * vc_restrict2(p, d1, d2, cx, cy) + half * vc_restrict2(p + d3, d1, d2, cx, cy)
            : vc_restrict2(p, d1, d2, cx, cy);
}

int orc_restrict_3d_vc_rf2(int cmask, const double *src, const int *spad, const int *sext, double *dst,
                           const int *dpad, const int *dext, const int *sb, const int *db, const int *srb,
                           const int *drb) {
  (void)srb;
  const box3 s = mkbox(sb), dbx = mkbox(db), r = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, s)) || (e = check_array(dpad, dext, dbx)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (r.str[d] <= 0)
      return CG_ERR_ARG;
    if (s.str[d] >= r.str[d] || dbx.str[d] != r.str[d])
      return CG_ERR_STRIDE;
    if (2 * s.str[d] != dbx.str[d])
      return CG_ERR_REFFACT;
    if (s.str[d] % 2)
      return CG_ERR_ALIGN;
  }
  if (empty(r))
    return CG_ERR_EMPTY;
  int n[3], so[3], dof[3], cen[3];
  for (int d = 0; d < 3; ++d) {
    cen[d] = (cmask >> d) & 1;
    if ((r.lo[d] - s.lo[d] - s.str[d] / 2) % s.str[d] || (r.lo[d] - dbx.lo[d]) % dbx.str[d])
      return CG_ERR_ALIGN;
    n[d] = shape(r, d) / r.str[d];
    so[d] = (r.lo[d] - s.lo[d] - s.str[d] / 2) / s.str[d];
    dof[d] = (r.lo[d] - dbx.lo[d]) / dbx.str[d];
    // fine points srcoff + 2i (+1 in a cell-centred direction) in the source
    if (so[d] < 0 || so[d] + 2 * (n[d] - 1) + cen[d] >= sext[d])
      return CG_ERR_CONTAIN;
  }
  if (!contained(r, dbx))
    return CG_ERR_CONTAIN;
  const long long d1 = 1, d2 = spad[0], d3 = (long long)spad[0] * spad[1];
#pragma omp parallel for collapse(2) if (g_use_openmp)
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i)
        dst[idx3(dof[0] + i, dof[1] + j, dof[2] + k, dpad)] =
            vc_restrict3(&src[idx3(so[0] + 2 * i, so[1] + 2 * j, so[2] + 2 * k, spad)], d1, d2, d3, cen[0],
                         cen[1], cen[2]);
  return CG_OK;
}

// Higher-order cell-centred restriction (use_higher_order_restriction,
// data.cc:1064-1094): restrict_3d_cc_o3_rf2.cc:33-212 evaluates at the coarse
// cell centre the cubic through 4 fine cells per axis, 64 terms with weights
// (w_i w_j w_k)/4096, w = {-1, 9, 9, -1}, listed with i fastest, then j,
// then k (:146-211); restrict_3d_cc_o5_rf2.cc:36-74 nests 6-point quintic
// sums, x inside y inside z, weights {3, -25, 150, 150, -25, 3}/256.  Both
// fit polynomials through the fine cell centres, so they are not
// conservative.  Canonical order: res = c0 * x0, then fma(c, x, res) term by
// term (o3 in the reference's listing order, o5 in its nesting).  The
// footprint -- fine cells 2i-1..2i+2 (o3) or 2i-2..2i+3 (o5) -- must lie in
// the source array (the reference's expand(1,1) check is one cell short of
// the o5 stencil).
int orc_restrict_3d_cc_ho_rf2(int order, const double *src, const int *spad, const int *sext,
                              double *dst, const int *dpad, const int *dext, const int *sb,
                              const int *db, const int *srb, const int *drb) {
  (void)srb;
  if (order != 3 && order != 5)
    return CG_ERR_ORDER;
  const box3 s = mkbox(sb), dbx = mkbox(db), r = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, s)) || (e = check_array(dpad, dext, dbx)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (r.str[d] <= 0)
      return CG_ERR_ARG;
    if (s.str[d] >= r.str[d] || dbx.str[d] != r.str[d])
      return CG_ERR_STRIDE;
    if (2 * s.str[d] != dbx.str[d])
      return CG_ERR_REFFACT;
    if (s.str[d] % 2)
      return CG_ERR_ALIGN;
  }
  if (empty(r))
    return CG_ERR_EMPTY;
  const int h = order == 3 ? 1 : 2;
  int n[3], so[3], dof[3];
  for (int d = 0; d < 3; ++d) {
    if ((r.lo[d] - s.lo[d] - s.str[d] / 2) % s.str[d] ||
        (r.lo[d] - dbx.lo[d]) % dbx.str[d])
      return CG_ERR_ALIGN;
    n[d] = shape(r, d) / r.str[d];
    so[d] = (r.lo[d] - s.lo[d] - s.str[d] / 2) / s.str[d];
    dof[d] = (r.lo[d] - dbx.lo[d]) / dbx.str[d];
    if (so[d] - h < 0 || so[d] + 2 * n[d] - 1 + h >= sext[d])
      return CG_ERR_CONTAIN;
  }
  if (!contained(r, dbx))
    return CG_ERR_CONTAIN;
#pragma omp parallel for collapse(2) if (g_use_openmp)
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        auto S = [&](int a, int b, int c) {
          return src[idx3(so[0] + 2 * i + a, so[1] + 2 * j + b, so[2] + 2 * k + c, spad)];
        };
        double res = 0;
        if (order == 3) {
          const double w[4] = {-1, 9, 9, -1};
          bool first = true;
          for (int c = 0; c < 4; ++c)
            for (int b = 0; b < 4; ++b)
              for (int a = 0; a < 4; ++a) {
                const double cf = w[a] * w[b] * w[c] / 4096;
                const double x = S(a - 1, b - 1, c - 1);
                res = first ? cf * x : std::fma(cf, x, res);
                first = false;
              }
        } else {
          const double w[6] = {3.0 / 256, -25.0 / 256, 150.0 / 256, 150.0 / 256, -25.0 / 256, 3.0 / 256};
          double r2[6];
          for (int c = 0; c < 6; ++c) {
            double r1[6];
            for (int b = 0; b < 6; ++b) {
              double x = w[0] * S(-2, b - 2, c - 2);
              for (int a = 1; a < 6; ++a)
                x = std::fma(w[a], S(a - 2, b - 2, c - 2), x);
              r1[b] = x;
            }
            double y = w[0] * r1[0];
            for (int b = 1; b < 6; ++b)
              y = std::fma(w[b], r1[b], y);
            r2[c] = y;
          }
          res = w[0] * r2[0];
          for (int c = 1; c < 6; ++c)
            res = std::fma(w[c], r2[c], res);
        }
        dst[idx3(dof[0] + i, dof[1] + j, dof[2] + k, dpad)] = res;
      }
  return CG_OK;
}

// Time-interpolation weights, interpolate_3d_2tl.cc:97-107 and
// interpolate_3d_3tl.cc:98-115; ntl in {1,2,3}.  Returns an error code for
// equal times or extrapolation, exactly as the reference checks.
int orc_time_factors(int ntl, const double *tt, double t, double *s) {
  const double eps = 1.0e-12;
  if (ntl == 1) {
    s[0] = 1.0;
    return CG_OK;
  }
  if (ntl == 2) {
    double t1 = tt[0], t2 = tt[1];
    if (std::fabs(t1 - t2) < eps)
      return CG_ERR_TIME_SAME;
    if (t < std::min(t1, t2) - eps || t > std::max(t1, t2) + eps)
      return CG_ERR_TIME_EXTRAP;
    s[0] = (t - t2) / (t1 - t2);
    s[1] = (t - t1) / (t2 - t1);
    return CG_OK;
  }
  if (ntl == 3) {
    double t1 = tt[0], t2 = tt[1], t3 = tt[2];
    if (std::fabs(t1 - t2) < eps || std::fabs(t1 - t3) < eps ||
        std::fabs(t2 - t3) < eps)
      return CG_ERR_TIME_SAME;
    if (t < std::fmin(std::fmin(t1, t2), t3) - eps ||
        t > std::fmax(std::fmax(t1, t2), t3) + eps)
      return CG_ERR_TIME_EXTRAP;
    s[0] = (t - t2) * (t - t3) / ((t1 - t2) * (t1 - t3));
    s[1] = (t - t1) * (t - t3) / ((t2 - t1) * (t2 - t3));
    s[2] = (t - t1) * (t - t2) / ((t3 - t1) * (t3 - t2));
    return CG_OK;
  }
  if (ntl == 4 || ntl == 5) {
    // interpolate_3d_4tl.cc:98-119, interpolate_3d_5tl.cc:101-128: every
    // pair apart by eps, no extrapolation, and per level the product of
    // (t - t_m) over the other levels in level order divided by the
    // product of (t_k - t_m) in the same order
    double tmin = tt[0], tmax = tt[0];
    for (int a = 0; a < ntl; ++a) {
      for (int b = a + 1; b < ntl; ++b)
        if (std::fabs(tt[a] - tt[b]) < eps)
          return CG_ERR_TIME_SAME;
      tmin = std::min(tmin, tt[a]);
      tmax = std::max(tmax, tt[a]);
    }
    if (t < tmin - eps || t > tmax + eps)
      return CG_ERR_TIME_EXTRAP;
    for (int k = 0; k < ntl; ++k) {
      double num = 0.0, den = 0.0;
      bool first = true;
      for (int m = 0; m < ntl; ++m) {
        if (m == k)
          continue;
        num = first ? (t - tt[m]) : num * (t - tt[m]);
        den = first ? (tt[k] - tt[m]) : den * (tt[k] - tt[m]);
        first = false;
      }
      s[k] = num / den;
    }
    return CG_OK;
  }
  return CG_ERR_ORDER;
}

// interpolate_3d_2tl / _3tl / _4tl / _5tl on same-stride arrays (ntl in
// 2..5); the sum in canonical order: s1*a, then fma(s2, b, .), ...
// (the reference writes s1fac*src1 + s2fac*src2 + ..., interpolate_3d_4tl.cc:125-128)
int orc_interpolate_3d(int ntl, const double *const *srcs, const double *tt,
                       const int *spad, const int *sext, double *dst, double t,
                       const int *dpad, const int *dext, const int *sb,
                       const int *db, const int *srb, const int *drb) {
  (void)srb;
  const box3 s = mkbox(sb), dbx = mkbox(db), r = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, s)) || (e = check_array(dpad, dext, dbx)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (r.str[d] <= 0)
      return CG_ERR_ARG;
    if (s.str[d] != r.str[d] || dbx.str[d] != r.str[d])
      return CG_ERR_STRIDE;
  }
  if (empty(r))
    return CG_ERR_EMPTY;
  if (!contained(r, s) || !contained(r, dbx))
    return CG_ERR_CONTAIN;
  double f[5];
  if (ntl > 5 || (e = orc_time_factors(ntl, tt, t, f)) || ntl < 2)
    return e ? e : CG_ERR_ORDER;
  int n[3], so[3], dof[3];
  for (int d = 0; d < 3; ++d) {
    if ((r.lo[d] - s.lo[d]) % s.str[d] || (r.lo[d] - dbx.lo[d]) % dbx.str[d])
      return CG_ERR_ALIGN;
    n[d] = sizes(r, d);
    so[d] = (r.lo[d] - s.lo[d]) / s.str[d];
    dof[d] = (r.lo[d] - dbx.lo[d]) / dbx.str[d];
  }
#pragma omp parallel for collapse(2) if (g_use_openmp)
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        size_t si = idx3(so[0] + i, so[1] + j, so[2] + k, spad);
        double res = f[0] * srcs[0][si];
        for (int l = 1; l < ntl; ++l)
          res = std::fma(f[l], srcs[l][si], res);
        dst[idx3(dof[0] + i, dof[1] + j, dof[2] + k, dpad)] = res;
      }
  return CG_OK;
}

// interpolate_eno_3d_3tl, interpolate_eno_3d_3tl.cc:24-164: quadratic
// interpolation through three time levels; where it makes a new extremum
// (above the largest or below the smallest of the three values) the point
// falls back to linear interpolation between the two levels whose interval
// holds t (levels 1-2 if t lies in theirs, else 2-3).  eps is relative to
// the time span (:107), unlike interpolate_3d_3tl's absolute 1e-12.
// Weights: the reference's expressions (:119-129); sums in canonical order.
int orc_eno_time_factors(const double *tt, double t, double *s3, double *s2, int *l0) {
  const double t1 = tt[0], t2 = tt[1], t3 = tt[2];
  const double tmin = std::fmin(std::fmin(t1, t2), t3);
  const double tmax = std::fmax(std::fmax(t1, t2), t3);
  const double eps = 1.0e-12 * (tmax - tmin);
  if (std::fabs(t1 - t2) < eps || std::fabs(t1 - t3) < eps || std::fabs(t2 - t3) < eps)
    return CG_ERR_TIME_SAME;
  if (t < tmin - eps || t > tmax + eps)
    return CG_ERR_TIME_EXTRAP;
  s3[0] = (t - t2) * (t - t3) / ((t1 - t2) * (t1 - t3));
  s3[1] = (t - t1) * (t - t3) / ((t2 - t1) * (t2 - t3));
  s3[2] = (t - t1) * (t - t2) / ((t3 - t1) * (t3 - t2));
  const bool use_12 = t >= std::min(t1, t2) - eps && t <= std::max(t1, t2) + eps;
  const bool use_23 = t >= std::min(t2, t3) - eps && t <= std::max(t2, t3) + eps;
  if (!use_12 && !use_23)
    return CG_ERR_TIME_EXTRAP;
  if (use_12) {
    *l0 = 0;
    s2[0] = (t - t2) / (t1 - t2);
    s2[1] = (t - t1) / (t2 - t1);
  } else {
    *l0 = 1;
    s2[0] = (t - t3) / (t2 - t3);
    s2[1] = (t - t2) / (t3 - t2);
  }
  return CG_OK;
}

static inline double eno_time_point(double a, double b, double c, const double *s3,
                                    const double *s2, int l0) {
  double d = std::fma(s3[2], c, std::fma(s3[1], b, s3[0] * a));
  if (d > std::max(std::max(a, b), c) || d < std::min(std::min(a, b), c)) {
    const double u = l0 == 0 ? a : b, v = l0 == 0 ? b : c;
    d = std::fma(s2[1], v, s2[0] * u);
  }
  return d;
}

int orc_interpolate_eno_3d_3tl(const double *const *srcs, const double *tt, const int *spad,
                               const int *sext, double *dst, double t, const int *dpad,
                               const int *dext, const int *sb, const int *db, const int *srb,
                               const int *drb) {
  (void)srb;
  const box3 s = mkbox(sb), dbx = mkbox(db), r = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, s)) || (e = check_array(dpad, dext, dbx)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (r.str[d] <= 0)
      return CG_ERR_ARG;
    if (s.str[d] != r.str[d] || dbx.str[d] != r.str[d])
      return CG_ERR_STRIDE;
  }
  if (empty(r))
    return CG_ERR_EMPTY;
  if (!contained(r, s) || !contained(r, dbx))
    return CG_ERR_CONTAIN;
  double s3[3], s2[2];
  int l0;
  if ((e = orc_eno_time_factors(tt, t, s3, s2, &l0)))
    return e;
  int n[3], so[3], dof[3];
  for (int d = 0; d < 3; ++d) {
    if ((r.lo[d] - s.lo[d]) % s.str[d] || (r.lo[d] - dbx.lo[d]) % dbx.str[d])
      return CG_ERR_ALIGN;
    n[d] = sizes(r, d);
    so[d] = (r.lo[d] - s.lo[d]) / s.str[d];
    dof[d] = (r.lo[d] - dbx.lo[d]) / dbx.str[d];
  }
#pragma omp parallel for collapse(2) if (g_use_openmp)
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        size_t si = idx3(so[0] + i, so[1] + j, so[2] + k, spad);
        dst[idx3(dof[0] + i, dof[1] + j, dof[2] + k, dpad)] =
            eno_time_point(srcs[0][si], srcs[1][si], srcs[2][si], s3, s2, l0);
      }
  return CG_OK;
}

// gdata::find_source_timelevel, gdata.cc:288-349.  op uses CG_OP_*.
int orc_find_source_timelevel(const double *times, int ntimes, double time,
                              int order_time, int op, int *timelevel0,
                              int *ntimelevels) {
  if (ntimes <= 0 || order_time < 0)
    return CG_ERR_ARG;
  const double eps = 1.0e-12;
  double mn = times[0], mx = times[0];
  for (int i = 1; i < ntimes; ++i) {
    mn = std::min(mn, times[i]);
    mx = std::max(mx, times[i]);
  }
  double some_time = std::fabs(mn) + std::fabs(mx) + 1.0;
  if (op != CG_OP_COPY)
    if (time < mn - eps * some_time || time > mx + eps * some_time)
      return CG_ERR_TIME_EXTRAP;
  int tl = -1;
  if (ntimes == 1)
    tl = 0;
  if (tl == -1 && op == CG_OP_COPY)
    tl = 0;
  if (tl == -1)
    for (int i = 0; i < ntimes; ++i)
      if (std::fabs(times[i] - time) < eps * some_time) {
        tl = i;
        break;
      }
  if (tl >= 0) {
    *timelevel0 = tl;
    *ntimelevels = 1;
  } else {
    *timelevel0 = 0;
    *ntimelevels = order_time + 1;
  }
  return CG_OK;
}

// data<T>::transfer_time + transfer_prolongate for the operators in scope,
// in the reference's shape: each needed time level is prolongated into a

// ---- staggered operators --------------------------------------------------
// prolongate_3d_stagger{011,101,110,111}.cc (the four differ only in which
// directions are staggered) with coeffs_prolongate_3d_stagger.hh, and
// restrict_3d_stagger{011,101,110,111}.cc with coeffs_restrict_3d_stagger.hh.
// stag: bit d set = staggered in direction d (011 -> 6, 101 -> 5, 110 -> 3,
// 111 -> 7).  The reference's loops are followed as written: per fine point
// dst = 0, then dst += coeff * coeff * coeff * src with the x taps outermost,
// and for ORDER 2 at a point coincident with a coarse vertex in the
// unstaggered direction the two-dimensional sum over the other two.

static int g_stagger = 7; // directions used by orc_prolongate_time's kind 6

void orc_set_stagger(int stag) { g_stagger = stag; }

// coeff[1..4][.] of coeffs_prolongate_3d_stagger.hh for ORDER 2 and 3
// (index 0 of the second subscript is tap 1 of the reference)
static const double stag_c1_o2[3] = {5.0 / 32.0, 15.0 / 16.0, -3.0 / 32.0};
static const double stag_c2_o2[3] = {-3.0 / 32.0, 15.0 / 16.0, 5.0 / 32.0};
static const double stag_c3_o2[3] = {-1.0 / 8.0, 3.0 / 4.0, 3.0 / 8.0};
static const double stag_c1_o3[4] = {-5.0 / 128.0, 35.0 / 128.0, 105.0 / 128.0, -7.0 / 128.0};
static const double stag_c2_o3[4] = {-7.0 / 128.0, 105.0 / 128.0, 35.0 / 128.0, -5.0 / 128.0};
static const double stag_c3_o3[4] = {-5.0 / 64.0, 37.0 / 64.0, 37.0 / 64.0, -5.0 / 64.0};
static const double stag_c4_o3[3] = {-1.0 / 32.0, 34.0 / 32.0, -1.0 / 32.0};

double orc_stagger_coeff(int order, int which, int tap) {
  if (order == 2) {
    const double *c = which == 1 ? stag_c1_o2 : which == 2 ? stag_c2_o2 : which == 3 ? stag_c3_o2 : nullptr;
    return c && tap >= 0 && tap < 3 ? c[tap] : 0.0;
  }
  if (order == 3) {
    if (which == 4)
      return tap >= 0 && tap < 3 ? stag_c4_o3[tap] : 0.0;
    const double *c = which == 1 ? stag_c1_o3 : which == 2 ? stag_c2_o3 : which == 3 ? stag_c3_o3 : nullptr;
    return c && tap >= 0 && tap < 4 ? c[tap] : 0.0;
  }
  return 0.0;
}

int orc_prolongate_3d_stagger(int stag, int order, const double *src, const int *spad, const int *sext,
                              double *dst, const int *dpad, const int *dext, const int *sb, const int *db,
                              const int *srb, const int *drb) {
  (void)srb;
  if (stag < 1 || stag > 7)
    return CG_ERR_ARG;
  if (order != 2 && order != 3)
    return CG_ERR_ORDER;
  const box3 srcbbox = mkbox(sb), dstbbox = mkbox(db), reg = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, srcbbox)) || (e = check_array(dpad, dext, dstbbox)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (reg.str[d] <= 0)
      return CG_ERR_ARG;
    if (srcbbox.str[d] <= reg.str[d] || dstbbox.str[d] != reg.str[d])
      return CG_ERR_STRIDE;
    if (srcbbox.str[d] != 2 * dstbbox.str[d])
      return CG_ERR_REFFACT;
    if (srcbbox.lo[d] % srcbbox.str[d] || dstbbox.lo[d] % dstbbox.str[d] || reg.lo[d] % reg.str[d])
      return CG_ERR_ALIGN;
  }
  if (empty(reg))
    return CG_ERR_EMPTY;
  if (!contained(reg, dstbbox))
    return CG_ERR_CONTAIN;
  int off[3], doff[3], n[3];
  for (int d = 0; d < 3; ++d) {
    const int num = reg.lo[d] - srcbbox.lo[d];
    if (num % reg.str[d] || (reg.lo[d] - dstbbox.lo[d]) % reg.str[d])
      return CG_ERR_ALIGN;
    off[d] = num / reg.str[d];
    doff[d] = (reg.lo[d] - dstbbox.lo[d]) / reg.str[d];
    n[d] = sizes(reg, d);
  }
  // per direction and fine index: first coarse tap, tap count, coefficients
  // (the index expressions of the reference's eight parity cases)
  struct Taps {
    int first, count;
    const double *c;
  };
  auto taps = [&](int d, int g) -> Taps {
    const int b = g >> 1;
    const bool st = (stag >> d) & 1, odd = g & 1;
    if (order == 2) {
      if (st)
        return {b - 1, 3, odd ? stag_c2_o2 : stag_c1_o2};
      return odd ? Taps{b - 1, 3, stag_c3_o2} : Taps{b, 1, nullptr};
    }
    if (st)
      return odd ? Taps{b - 1, 4, stag_c2_o3} : Taps{b - 2, 4, stag_c1_o3};
    return odd ? Taps{b - 1, 4, stag_c3_o3} : Taps{b - 1, 3, stag_c4_o3};
  };
  // every tap of every point must lie in the source array
  for (int d = 0; d < 3; ++d)
    for (int f = 0; f < n[d]; ++f) {
      const Taps t = taps(d, off[d] + f);
      if (off[d] < 0 || t.first < 0 || t.first + t.count > sext[d])
        return CG_ERR_CONTAIN;
    }
  auto body = [&](const box3 &r) {
    for (int k = r.lo[2]; k <= r.up[2]; ++k)
      for (int j = r.lo[1]; j <= r.up[1]; ++j)
        for (int i = r.lo[0]; i <= r.up[0]; ++i) {
          const Taps tx = taps(0, off[0] + i), ty = taps(1, off[1] + j), tz = taps(2, off[2] + k);
          double acc = 0.0;
          if (!tx.c) { // ORDER 2, coincident in x (unstaggered): y outer, z inner
            for (int b = 0; b < ty.count; ++b)
              for (int c = 0; c < tz.count; ++c)
                acc += ty.c[b] * tz.c[c] * src[idx3(tx.first, ty.first + b, tz.first + c, spad)];
          } else if (!ty.c) { // coincident in y: x outer, z inner
            for (int a = 0; a < tx.count; ++a)
              for (int c = 0; c < tz.count; ++c)
                acc += tx.c[a] * tz.c[c] * src[idx3(tx.first + a, ty.first, tz.first + c, spad)];
          } else if (!tz.c) { // coincident in z: x outer, y inner
            for (int a = 0; a < tx.count; ++a)
              for (int b = 0; b < ty.count; ++b)
                acc += tx.c[a] * ty.c[b] * src[idx3(tx.first + a, ty.first + b, tz.first, spad)];
          } else {
            for (int a = 0; a < tx.count; ++a)
              for (int b = 0; b < ty.count; ++b)
                for (int c = 0; c < tz.count; ++c)
                  acc += tx.c[a] * ty.c[b] * tz.c[c] *
                         src[idx3(tx.first + a, ty.first + b, tz.first + c, spad)];
          }
          dst[idx3(doff[0] + i, doff[1] + j, doff[2] + k, dpad)] = acc;
        }
  };
  box3 all{{0, 0, 0}, {n[0] - 1, n[1] - 1, n[2] - 1}, {1, 1, 1}};
  if (g_use_openmp)
    split_longest(all, 0, body);
  else
    body(all);
  return CG_OK;
}

int orc_restrict_3d_stagger(int stag, const double *src, const int *spad, const int *sext, double *dst,
                            const int *dpad, const int *dext, const int *sb, const int *db, const int *srb,
                            const int *drb) {
  (void)srb;
  if (stag < 1 || stag > 7)
    return CG_ERR_ARG;
  const box3 srcbbox = mkbox(sb), dstbbox = mkbox(db), reg = mkbox(drb);
  int e;
  if ((e = check_array(spad, sext, srcbbox)) || (e = check_array(dpad, dext, dstbbox)))
    return e;
  for (int d = 0; d < 3; ++d) {
    if (reg.str[d] <= 0)
      return CG_ERR_ARG;
    if (srcbbox.str[d] >= reg.str[d] || dstbbox.str[d] != reg.str[d])
      return CG_ERR_STRIDE;
    if (2 * srcbbox.str[d] != dstbbox.str[d])
      return CG_ERR_REFFACT;
  }
  if (empty(reg))
    return CG_ERR_EMPTY;
  if (!contained(reg, dstbbox))
    return CG_ERR_CONTAIN;
  // coeff / coeff_i of coeffs_restrict_3d_stagger.hh (ORDER_STAG 3)
  static const double cs[4] = {-1.0 / 16.0, 9.0 / 16.0, 9.0 / 16.0, -1.0 / 16.0};
  static const double cv[3] = {1.0 / 8.0, 6.0 / 8.0, 1.0 / 8.0};
  int so[3], doff[3], n[3], cnt[3];
  const double *cf[3];
  for (int d = 0; d < 3; ++d) {
    if ((reg.lo[d] - srcbbox.lo[d]) % srcbbox.str[d] || (reg.lo[d] - dstbbox.lo[d]) % dstbbox.str[d])
      return CG_ERR_ALIGN;
    so[d] = (reg.lo[d] - srcbbox.lo[d]) / srcbbox.str[d];
    doff[d] = (reg.lo[d] - dstbbox.lo[d]) / dstbbox.str[d];
    n[d] = sizes(reg, d);
    const bool st = (stag >> d) & 1;
    cnt[d] = st ? 4 : 3;
    cf[d] = st ? cs : cv;
    if (so[d] - 1 < 0 || so[d] + 2 * (n[d] - 1) - 1 + cnt[d] > sext[d])
      return CG_ERR_CONTAIN;
  }
  for (int k = 0; k < n[2]; ++k)
    for (int j = 0; j < n[1]; ++j)
      for (int i = 0; i < n[0]; ++i) {
        double acc = 0.0;
        for (int a = 0; a < cnt[0]; ++a)
          for (int b = 0; b < cnt[1]; ++b)
            for (int c = 0; c < cnt[2]; ++c)
              acc += cf[0][a] * cf[1][b] * cf[2][c] *
                     src[idx3(so[0] + 2 * i - 1 + a, so[1] + 2 * j - 1 + b, so[2] + 2 * k - 1 + c, spad)];
        dst[idx3(doff[0] + i, doff[1] + j, doff[2] + k, dpad)] = acc;
      }
  return CG_OK;
}

// region-sized temporary (data.cc:504-515), then the temporaries are
// interpolated in time (data.cc:517, 1182-1299).  kind: 0 vertex Lagrange,
// 1 cell Lagrange, 2 cell ENO (order = ENO order), 3 cell TVD, 4 vertex ENO,
// 5 monotone order 5, 6 staggered (order 2 or 3, directions of
// orc_set_stagger).
int orc_prolongate_time(int kind, int order, const double *const *srcs,
                        const double *times, int ntimes, const int *spad,
                        const int *sext, double *dst, double time,
                        const int *dpad, const int *dext, const int *sb,
                        const int *db, const int *drb, int order_time) {
  int op = (kind == 2 || kind == 4) ? CG_OP_ENO : kind == 3 ? CG_OP_TVD : CG_OP_LAGRANGE;
  int tl0, ntl, e;
  if ((e = orc_find_source_timelevel(times, ntimes, time, order_time, op, &tl0, &ntl)))
    return e;
  if (ntl > ntimes)
    return CG_ERR_ARG;
  auto prolong = [&](const double *src, double *out, const int *opad,
                     const int *oext, const int *ob) -> int {
    switch (kind) {
    case 0:
      return orc_prolongate_3d_rf2(order, src, spad, sext, out, opad, oext, sb, ob, drb, drb);
    case 1:
      return orc_prolongate_3d_cc_rf2(order, src, spad, sext, out, opad, oext, sb, ob, drb, drb);
    case 2:
      return orc_prolongate_3d_cc_eno_rf2(order, src, spad, sext, out, opad, oext, sb, ob, drb, drb);
    case 3:
      return orc_prolongate_3d_cc_tvd(src, spad, sext, out, opad, oext, sb, ob, drb, drb);
    case 4:
      return orc_prolongate_3d_eno(src, spad, sext, out, opad, oext, sb, ob, drb, drb);
    case 5:
      return orc_prolongate_3d_o5_monotone_rf2(src, spad, sext, out, opad, oext, sb, ob, drb, drb);
    case 6:
      return orc_prolongate_3d_stagger(g_stagger, order, src, spad, sext, out, opad, oext, sb, ob, drb, drb);
    }
    return CG_ERR_OPERATOR;
  };
  if (ntl == 1)
    return prolong(srcs[tl0], dst, dpad, dext, db);
  // 4 and 5 levels: Lagrange-type groups only (data.cc:1192-1258; the
  // ENO/WENO/TVD/monotone branch asserts beyond order_time 2, :1260-1295)
  if (ntl > 5 || (ntl > 3 && kind != 0 && kind != 1 && kind != 6))
    return CG_ERR_ORDER;
  // temporaries with extent = region, unpadded (data.cc:507-509)
  const box3 r = mkbox(drb);
  int text[3];
  for (int d = 0; d < 3; ++d)
    text[d] = sizes(r, d);
  size_t npts = (size_t)text[0] * text[1] * text[2];
  std::vector<std::vector<double>> tmp(ntl, std::vector<double>(npts));
  for (int l = 0; l < ntl; ++l)
    if ((e = prolong(srcs[l], tmp[l].data(), text, text, drb)))
      return e;
  const double *tp[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  for (int l = 0; l < ntl; ++l)
    tp[l] = tmp[l].data();
  // ENO, WENO, TVD and monotone groups interpolate 3 time levels with
  // interpolate_eno_3d_3tl
  // (data.cc:1282-1292), Lagrange with interpolate_3d_2tl/3tl
  // (kinds 2 cell ENO, 3 TVD, 4 vertex ENO/WENO, 5 monotone: data.cc:1260-1292)
  if (ntl == 3 && (kind == 2 || kind == 3 || kind == 4 || kind == 5))
    return orc_interpolate_eno_3d_3tl(tp, times, text, text, dst, time, dpad, dext, drb, db,
                                      drb, drb);
  return orc_interpolate_3d(ntl, tp, times, text, text, dst, time, dpad, dext,
                            drb, db, drb, drb);
}

// ---- a level step's transfers on the CPU (the timed CPU baseline) ---------
// One transfer of a vertex-centred level step: 0 copy_3d (no slab),
// 1 restrict_3d_rf2, 2 prolongate_3d_rf2<order> with ntl source levels
// interpolated in time (ntl = 1: none).
struct orc_task {
  int op, order, ntl;
  const double *src[3];
  int spad[3], sext[3], sb[9];
  double *dst;
  int dpad[3], dext[3], db[9];
  int rb[9];
  double times[3], time;
};

// mode 0, the reference's shape: one operator call per transfer in order,
// each parallelised inside as the reference does it (call_operator's split
// along the longest axis for prolongation, collapse loops for copies and
// restriction) and time interpolation through region-sized temporaries
// (data.cc:56-100, 447-452, 497-521, 1100-1105).
// mode 1, the best this restatement does on a CPU: the transfers' planes
// as one work list under one parallel loop (no fork/join per region), each
// destination point of a prolongation evaluated on its ntl levels and
// combined at once (no temporaries); the same arithmetic, so the same bits.
// Returns 0 or the first failing operator's code.
int orc_replay(int mode, int n, const orc_task *t) {
  if (mode == 0) {
    for (int i = 0; i < n; ++i) {
      const orc_task &x = t[i];
      int e;
      switch (x.op) {
      case 0:
        e = orc_copy_3d(x.src[0], x.spad, x.sext, x.dst, x.dpad, x.dext, x.sb, x.db, x.rb, x.rb, nullptr);
        break;
      case 1:
        e = orc_restrict_3d_rf2(x.src[0], x.spad, x.sext, x.dst, x.dpad, x.dext, x.sb, x.db, x.rb, x.rb);
        break;
      case 2:
        e = orc_prolongate_time(0, x.order, x.src, x.times, x.ntl, x.spad, x.sext, x.dst, x.time, x.dpad,
                                x.dext, x.sb, x.db, x.rb, x.ntl - 1);
        break;
      default:
        e = CG_ERR_OPERATOR;
      }
      if (e)
        return e;
    }
    return CG_OK;
  }
  // fused: validate every transfer by its per-call operator first (on a
  // copy of nothing: the checks run before any store), then one work list
  struct Job {
    int task, k;
  };
  std::vector<Job> jobs;
  std::vector<std::array<int, 9>> geo((size_t)n); // n[3], source offset[3], destination offset[3]
  std::vector<std::array<double, 3>> w((size_t)n);
  for (int i = 0; i < n; ++i) {
    const orc_task &x = t[i];
    const box3 s = mkbox(x.sb), d = mkbox(x.db), r = mkbox(x.rb);
    int e;
    if ((e = check_array(x.spad, x.sext, s)) || (e = check_array(x.dpad, x.dext, d)))
      return e;
    if (empty(r))
      return CG_ERR_EMPTY;
    if (!contained(r, d))
      return CG_ERR_CONTAIN;
    for (int q = 0; q < 3; ++q) {
      const int m = sizes(r, q);
      geo[i][q] = m;
      geo[i][6 + q] = (r.lo[q] - d.lo[q]) / d.str[q];
      if (x.op == 2) {
        // fine index of the region start relative to the source array
        geo[i][3 + q] = (r.lo[q] - s.lo[q]) / r.str[q];
        const Stencil1D st = vc_stencil(x.order);
        const int g0 = geo[i][3 + q], g1 = g0 + m - 1;
        if ((g0 >> 1) + st.tmin[1] < 0 || (g1 >> 1) + st.tmin[1] + st.ntap[1] - 1 >= x.sext[q])
          return CG_ERR_CONTAIN;
      } else {
        geo[i][3 + q] = (r.lo[q] - s.lo[q]) / s.str[q];
        const int hi = geo[i][3 + q] + (x.op == 1 ? 2 * (m - 1) : m - 1);
        if (geo[i][3 + q] < 0 || hi >= x.sext[q])
          return CG_ERR_CONTAIN;
      }
    }
    w[i] = {1.0, 0.0, 0.0};
    if (x.op == 2 && x.ntl > 1 && (e = orc_time_factors(x.ntl, x.times, x.time, w[i].data())))
      return e;
    for (int k = 0; k < geo[i][2]; ++k)
      jobs.push_back({i, k});
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (size_t jb = 0; jb < jobs.size(); ++jb) {
    const orc_task &x = t[jobs[jb].task];
    const std::array<int, 9> &g = geo[jobs[jb].task];
    const int k = jobs[jb].k;
    if (x.op == 2) {
      const Stencil1D st = vc_stencil(x.order);
      const std::array<double, 3> &s3 = w[jobs[jb].task];
      const int gk = g[5] + k;
      for (int j = 0; j < g[1]; ++j) {
        const int gj = g[4] + j;
        for (int i = 0; i < g[0]; ++i) {
          const int gi = g[3] + i;
          double v[3] = {0.0, 0.0, 0.0};
          for (int l = 0; l < x.ntl; ++l)
            v[l] = lag3(st, x.src[l], x.spad, gi >> 1, gj >> 1, gk >> 1, gi & 1, gj & 1, gk & 1);
          double out = v[0];
          if (x.ntl == 2)
            out = std::fma(s3[1], v[1], s3[0] * v[0]);
          else if (x.ntl == 3)
            out = std::fma(s3[2], v[2], std::fma(s3[1], v[1], s3[0] * v[0]));
          x.dst[idx3(g[6] + i, g[7] + j, g[8] + k, x.dpad)] = out;
        }
      }
    } else {
      const int f = x.op == 1 ? 2 : 1;
      for (int j = 0; j < g[1]; ++j)
        for (int i = 0; i < g[0]; ++i)
          x.dst[idx3(g[6] + i, g[7] + j, g[8] + k, x.dpad)] =
              x.src[0][idx3(g[3] + f * i, g[4] + f * j, g[5] + f * k, x.spad)];
    }
  }
  return CG_OK;
}

} // extern "C"
