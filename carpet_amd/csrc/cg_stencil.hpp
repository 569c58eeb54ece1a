// cg_stencil.hpp -- 1-D line operators of the tensor-product prolongations.
//
// Every prolongation in scope except TVD is a nesting x -> y -> z of one 1-D
// operator: the value at a fine point of parity p (0 or 1 along an axis)
// is computed from a window of coarse values around the coarse base index
// b = g >> 1 (g = fine index relative to the source array in fine units).
// A window is described per parity by its first offset TMIN[p] and width
// NTAP[p]; a vertex-centred p = 0 is a pass-through.  Because the reference
// evaluates interp3 = op_z(interp2(...)), interp2 = op_y(interp1(...)),
// interp1 = op_x(values) and each op only sees its own line, computing all
// x-results once, then all y-results, then the z-results performs the same
// floating-point operations in the same order.
//
// Sources: CarpetLib/src/prolongate_3d_rf2.cc:27-124, 158-295 (vertex
// Lagrange), CarpetLib/src/prolongate_3d_cc_rf2.cc:49-195 (cell Lagrange),
// CarpetLib/src/prolongate_3d_cc_eno_rf2.cc:51-881 (cell ENO),
// CarpetLib/src/prolongate_3d_real8_eno.F90:8-67, 223-288 (vertex ENO).
#pragma once

#include <hip/hip_runtime.h>

namespace cg {

#define CG_HD __host__ __device__ __forceinline__

// LK_VENO is 4: kinds share numbers with the prolongation kinds of the
// planner, where 3 is the (point-wise) cell TVD
enum LineKind { LK_VLAG = 0, LK_CLAG = 1, LK_ENO = 2, LK_VENO = 4 };
// line kinds that exist in FP64 only (no FP32-storage modes)
constexpr bool fp64_only(int kind) { return kind == LK_ENO || kind == LK_VENO; }

// ---- weights -------------------------------------------------------------
template <int O> struct VW;
template <> struct VW<1> { static constexpr double c[2] = {1 / 2.0, 1 / 2.0}; };
template <> struct VW<3> {
  static constexpr double c[4] = {-1 / 16.0, 9 / 16.0, 9 / 16.0, -1 / 16.0};
};
template <> struct VW<5> {
  static constexpr double c[6] = {3 / 256.0,   -25 / 256.0, 150 / 256.0,
                                  150 / 256.0, -25 / 256.0, 3 / 256.0};
};
template <> struct VW<7> {
  static constexpr double c[8] = {-5 / 2048.0,   49 / 2048.0,  -245 / 2048.0,
                                  1225 / 2048.0, 1225 / 2048.0, -245 / 2048.0,
                                  49 / 2048.0,   -5 / 2048.0};
};
template <> struct VW<9> {
  static constexpr double c[10] = {
      35 / 65536.0,    -405 / 65536.0, 2268 / 65536.0, -8820 / 65536.0,
      39690 / 65536.0, 39690 / 65536.0, -8820 / 65536.0, 2268 / 65536.0,
      -405 / 65536.0,  35 / 65536.0};
};
template <> struct VW<11> {
  static constexpr double c[12] = {
      -63 / 524288.0,     847 / 524288.0,    -5445 / 524288.0,
      22869 / 524288.0,   -76230 / 524288.0, 320166 / 524288.0,
      320166 / 524288.0,  -76230 / 524288.0, 22869 / 524288.0,
      -5445 / 524288.0,   847 / 524288.0,    -63 / 524288.0};
};

template <int O> struct CW;
template <> struct CW<0> { static constexpr double c[1] = {1 / 1.0}; };
template <> struct CW<1> { static constexpr double c[2] = {1 / 4.0, 3 / 4.0}; };
template <> struct CW<2> {
  static constexpr double c[3] = {5 / 32.0, 30 / 32.0, -3 / 32.0};
};
template <> struct CW<3> {
  static constexpr double c[4] = {-5 / 128.0, 35 / 128.0, 105 / 128.0,
                                  -7 / 128.0};
};
template <> struct CW<4> {
  static constexpr double c[5] = {-45 / 2048.0, 420 / 2048.0, 1890 / 2048.0,
                                  -252 / 2048.0, 35 / 2048.0};
};
template <> struct CW<5> {
  static constexpr double c[6] = {63 / 8192.0,   -495 / 8192.0, 2310 / 8192.0,
                                  6930 / 8192.0, -693 / 8192.0, 77 / 8192.0};
};

// fused multiply-add in the canonical order
CG_HD double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Compensated FP32 dot product sum_t c(t)*v(t) (Ogita, Rump and Oishi,
// "Accurate sum and dot product", SIAM J. Sci. Comput. 26 (2005), Dot2):
// every product is split exactly with an FMA (TwoProduct) and every partial
// sum with Knuth's TwoSum; the errors are summed separately and added at the
// end, so the result is as accurate as if computed in twice the working
// precision and then rounded to FP32.
template <int N, class C, class V> CG_HD float dot2f(C c, V v) {
  float p = c(0) * v(0);
  float s = __builtin_fmaf(c(0), v(0), -p);
#pragma unroll
  for (int t = 1; t < N; ++t) {
    const float h = c(t) * v(t);
    const float r = __builtin_fmaf(c(t), v(t), -h);
    const float q = p + h;
    const float z = q - p;
    const float e = (p - (q - z)) + (h - z);
    p = q;
    s = s + (e + r);
  }
  return p + s;
}

// ---- line operator traits ------------------------------------------------
// Members: TMIN[2], NTAP[2], TMINALL, TMAXALL, R (widest window),
// template <int P, class V> static double apply(V v): v(t) gives the t-th
// value of the window of parity P (t = 0..NTAP[P]-1).

template <int KIND, int O> struct Line;

// vertex-centred Lagrange (coeffs1d<RT,ORDER>: imin = -(O+1)/2 + 1)
template <int O> struct Line<LK_VLAG, O> {
  static constexpr int N = O + 1;
  static constexpr int TMIN[2] = {0, -N / 2 + 1};
  static constexpr int NTAP[2] = {1, N};
  static constexpr int TMINALL = -N / 2 + 1;
  static constexpr int TMAXALL = N / 2;
  static constexpr int R = N;
  template <int P, class V> static CG_HD double apply(V v) {
    if constexpr (P == 0) {
      return v(0);
    } else {
      double r = 0.0;
#pragma unroll
      for (int t = 0; t < N; ++t)
        r = fmad(VW<O>::c[t], v(t), r);
      return r;
    }
  }
  // mixed precision: FP32 values, compensated FP32 accumulation (the weights
  // are dyadic fractions, exact in FP32)
  template <int P, class V> static CG_HD float apply_f(V v) {
    if constexpr (P == 0)
      return v(0);
    else
      return dot2f<N>([](int t) { return (float)VW<O>::c[t]; }, v);
  }
};

// cell-centred Lagrange (coeffs1d<RT,ORDER,di>: imin = -(O+1)/2 + (O odd ?
// di : 0); di == 1 mirrors the weights)
template <int O> struct Line<LK_CLAG, O> {
  static constexpr int N = O + 1;
  static constexpr int TMIN[2] = {-N / 2, -N / 2 + (O % 2 != 0 ? 1 : 0)};
  static constexpr int NTAP[2] = {N, N};
  static constexpr int TMINALL = -N / 2;
  static constexpr int TMAXALL = -N / 2 + (O % 2 != 0 ? 1 : 0) + N - 1;
  static constexpr int R = N;
  template <int P, class V> static CG_HD double apply(V v) {
    double r = 0.0;
#pragma unroll
    for (int t = 0; t < N; ++t)
      r = fmad(CW<O>::c[P == 0 ? t : N - 1 - t], v(t), r);
    return r;
  }
  template <int P, class V> static CG_HD float apply_f(V v) {
    return dot2f<N>([](int t) { return (float)CW<O>::c[P == 0 ? t : N - 1 - t]; }, v);
  }
};

// cell-centred ENO, orders 2 and 3 (coeffs1d<RT,ORDER,di,OFFSET>).  The
// window of parity di is [minimin, maximax) of coeffs1d<RT,ORDER,di,0>:
//   ENO<2>: di=0 -> [-2,2), di=1 -> [-1,3);  ENO<3>: [di-3, di+3).
template <int O> struct EnoW;
template <> struct EnoW<2> {
  static constexpr int MM[2] = {-2, -1};
  static constexpr int W = 4;
};
template <> struct EnoW<3> {
  static constexpr int MM[2] = {-3, -2};
  static constexpr int W = 6;
};

template <int O> struct Line<LK_ENO, O> {
  static constexpr int TMIN[2] = {EnoW<O>::MM[0], EnoW<O>::MM[1]};
  static constexpr int NTAP[2] = {EnoW<O>::W, EnoW<O>::W};
  static constexpr int TMINALL = EnoW<O>::MM[0];
  static constexpr int TMAXALL = EnoW<O>::MM[1] + EnoW<O>::W - 1;
  static constexpr int R = EnoW<O>::W;

  // f[i - MM] for coarse offsets i; the linear fallback uses
  // coeffs1d<RT,1,di,0> = {1/4, 3/4} at imin = -1 + di (mirrored for di=1)
  template <int P, class V> static CG_HD double linear(V v) {
    constexpr int mm = EnoW<O>::MM[P];
    double r = 0.0;
    if constexpr (P == 0) {
      r = fmad(1 / 4.0, v(-1 - mm), r);
      r = fmad(3 / 4.0, v(0 - mm), r);
    } else {
      r = fmad(3 / 4.0, v(0 - mm), r);
      r = fmad(1 / 4.0, v(1 - mm), r);
    }
    return r;
  }

  // ENO stays FP64 in the mixed configuration
  template <int P, class V> static CG_HD float apply_f(V v) { return (float)apply<P>(v); }

  template <int P, class V> static CG_HD double apply(V v) {
    constexpr int mm = EnoW<O>::MM[P];
    double res = 0.0;
    if constexpr (O == 2) {
      // rcoeffs: OFFSET 0, imin = -1; lcoeffs: OFFSET (P==0 ? -1 : +1)
      // weights coeffs1dc<2,0> = {5,30,-3}/32, coeffs1dc<2,-1> =
      // {-0.09375,0.4375,0.65625}, mirrored for P == 1
      constexpr int limin = P == 0 ? -2 : 0;
      constexpr int rimin = -1;
      double lV = 0.0, rV = 0.0;
      lV = fmad(+1.0, v(limin + 0 - mm), lV);
      lV = fmad(-2.0, v(limin + 1 - mm), lV);
      lV = fmad(+1.0, v(limin + 2 - mm), lV);
      rV = fmad(+1.0, v(rimin + 0 - mm), rV);
      rV = fmad(-2.0, v(rimin + 1 - mm), rV);
      rV = fmad(+1.0, v(rimin + 2 - mm), rV);
      if (lV * rV <= 0) {
        res = linear<P>(v);
      } else if (__builtin_fabs(lV) < __builtin_fabs(rV)) {
        constexpr double L[3] = {-0.09375, +0.4375, +0.65625};
#pragma unroll
        for (int t = 0; t < 3; ++t)
          res = fmad(L[P == 0 ? t : 2 - t], v(limin + t - mm), res);
      } else {
        constexpr double C[3] = {5 / 32.0, 30 / 32.0, -3 / 32.0};
#pragma unroll
        for (int t = 0; t < 3; ++t)
          res = fmad(C[P == 0 ? t : 2 - t], v(rimin + t - mm), res);
      }
    } else {
      // O == 3: lcoeffs OFFSET -1, ccoeffs 0, rcoeffs +1; imin = -2 + P + OFFSET
      // weights coeffs1dc<3,OFF'> with OFF' = (P==0 ? OFFSET : -OFFSET),
      // mirrored for P == 1; differences {1,-3,3,-1} are not mirrored
      constexpr int i0 = -2 + P;
      constexpr double D[4] = {+1.0, -3.0, +3.0, -1.0};
      constexpr double W0[4] = {-0.0390625, 0.2734375, 0.8203125, -0.0546875};
      constexpr double Wp[4] = {0.1171875, 1.0546875, -0.2109375, 0.0390625};
      constexpr double Wm[4] = {0.0546875, -0.2578125, 0.6015625, 0.6015625};
      double V0 = 0.0, V1 = 0.0, V2 = 0.0;
#pragma unroll
      for (int t = 0; t < 4; ++t)
        V0 = fmad(D[t], v(i0 - 1 + t - mm), V0);
#pragma unroll
      for (int t = 0; t < 4; ++t)
        V1 = fmad(D[t], v(i0 + t - mm), V1);
#pragma unroll
      for (int t = 0; t < 4; ++t)
        V2 = fmad(D[t], v(i0 + 1 + t - mm), V2);
      if (V0 * V2 <= 0 || V0 * V1 <= 0 || V1 * V2 <= 0)
        return linear<P>(v);
      int mn = 1;
      if (__builtin_fabs(V0) < __builtin_fabs(V1))
        mn = 0;
      if (__builtin_fabs(V2) < __builtin_fabs(mn == 0 ? V0 : V1))
        mn = 2;
      if (mn == 0) {
        // OFFSET -1 -> table (P==0 ? -1 : +1)
#pragma unroll
        for (int t = 0; t < 4; ++t)
          res = fmad(P == 0 ? Wm[t] : Wp[3 - t], v(i0 - 1 + t - mm), res);
      } else if (mn == 1) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
          res = fmad(W0[P == 0 ? t : 3 - t], v(i0 + t - mm), res);
      } else {
        // OFFSET +1 -> table (P==0 ? +1 : -1)
#pragma unroll
        for (int t = 0; t < 4; ++t)
          res = fmad(P == 0 ? Wp[t] : Wm[3 - t], v(i0 + 1 + t - mm), res);
      }
    }
    // monotonicity check against f[-mm-1+di] and f[-mm+di]
    // (prolongate_3d_cc_eno_rf2.cc:378-408)
    const double fa = v(-1 + P - mm), fb = v(P - mm);
    if ((res - fa) * (fb - res) < 0)
      res = linear<P>(v);
    return res;
  }
};

// vertex-centred ENO of the midpoint between q2 and q3
// (prolongate_3d_real8_eno.F90:8-67): the quadratic through q1..q3 or
// q2..q4, whichever has the smaller second difference (the right one on a
// tie), evaluated at the midpoint; the average of q2 and q3 where that
// value leaves [q2, q3] or the second differences differ in sign or vanish.
// Evaluated as written, every product and sum rounded on its own.
CG_HD double eno1d_v(double q1, double q2, double q3, double q4) {
  const double dl = (q1 + q3) - 2.0 * q2;
  const double dr = (q2 + q4) - 2.0 * q3;
  double e;
  if (__builtin_fabs(dl) < __builtin_fabs(dr))
    e = 0.125 * ((-q1 + 6.0 * q2) + 3.0 * q3);
  else
    e = 0.125 * ((3.0 * q2 + 6.0 * q3) - q4);
  if ((e - q2) * (q3 - e) < 0.0 || dl * dr <= 0.0)
    e = 0.5 * (q2 + q3);
  return e;
}

// vertex-centred ENO (prolongate_3d_real8_eno.F90:119-282): a coincident
// child passes the coarse value through, a midpoint child is eno1d of the
// coarse points b-1..b+2; one scheme (the dispatch maps ENO order_space 3
// and 5 and WENO order_space 5 onto it, data.cc:654-762)
template <int O> struct Line<LK_VENO, O> {
  static constexpr int TMIN[2] = {0, -1};
  static constexpr int NTAP[2] = {1, 4};
  static constexpr int TMINALL = -1;
  static constexpr int TMAXALL = 2;
  static constexpr int R = 4;
  template <int P, class V> static CG_HD double apply(V v) {
    if constexpr (P == 0)
      return v(0);
    else
      return eno1d_v(v(0), v(1), v(2), v(3));
  }
  template <int P, class V> static CG_HD float apply_f(V v) { return (float)apply<P>(v); }
};

// host-side view of a line operator's windows (for validation and
// footprints); kind/order are runtime values there
struct LineInfo {
  int tmin[2];
  int ntap[2];
  int tminall, tmaxall;
};

template <int KIND, int O> LineInfo line_info() {
  using L = Line<KIND, O>;
  LineInfo li;
  for (int p = 0; p < 2; ++p) {
    li.tmin[p] = L::TMIN[p];
    li.ntap[p] = L::NTAP[p];
  }
  li.tminall = L::TMINALL;
  li.tmaxall = L::TMAXALL;
  return li;
}

} // namespace cg
