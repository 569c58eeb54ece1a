// cg_tiles.hpp -- device bodies of the transfer kernels, shared by the
// per-call launches (cg_kernels.hip) and the persistent batch kernels
// (cg_batch_kernels.hip).
//
// k_tensor: every tensor-product prolongation (vertex Lagrange
// prolongate_3d_rf2, cell Lagrange prolongate_3d_cc_rf2, cell ENO
// prolongate_3d_cc_eno_rf2), optionally fused with 2- or 3-level time
// interpolation.  A workgroup owns a tile of CX x CY coarse bases in x-y
// (a 2CX x 2CY fine column) and marches through a chunk of coarse z planes:
//   1. the coarse plane window (tile + stencil halo, clipped to the region's
//      exact source footprint) is loaded into LDS with coalesced x-rows;
//   2. x-pass: every coarse row of the window is interpolated to both fine
//      x children of each base (LDS -> LDS);
//   3. y-pass: every fine x column is interpolated to both fine y children
//      of each base; the results stay in registers in a ring of the last R
//      coarse planes;
//   4. z-emission: as soon as the ring holds all planes of a fine z point,
//      that fine plane is combined over time levels and stored (one
//      coalesced x-row per wave).
// Each pass evaluates the reference's nested sums on the same operands in
// the same order, so the result equals the per-point nested evaluation bit
// for bit while the work per fine point drops from (O+1)^3 to ~3(O+1)/2.
// The kernel is bandwidth-bound: 8 bytes written per fine point plus ~1/8 of
// that read per source time level.
//
// k_point: copy_3d, restrict_3d_rf2, restrict_3d_cc_rf2,
// interpolate_3d_2tl/3tl and prolongate_3d_cc_tvd, one destination point
// per thread over a flattened x-y plane (so thin ghost slabs still fill
// waves) and planes along blockIdx.y.

#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "cg_kernels.hpp"
#include "cg_stencil.hpp"

namespace cg {


// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; remap
// the hardware index so that each XCD walks a contiguous range of tiles and
// neighbouring tiles share stencil halos through the same L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  constexpr int NX = 8;
  if (n < NX)
    return b;
  const int per = n / NX, rem = n % NX;
  const int xcd = b % NX, slot = b / NX;
  return xcd * per + min(xcd, rem) + slot;
}

template <typename TS> __device__ __forceinline__ double ld(const void *p, long long o) {
  return (double)(static_cast<const TS *>(p)[o]);
}

// LDS footprint (doubles) of one tensor tile; XQ = bases per thread in the
// x-pass (2: rows padded to an even stride for 16-byte LDS accesses); NB =
// buffers of the window and of the x-pass results (2 for the pipelined
// schedule)
template <class L, int NTL, int CX, int CY, int XQ = 1, int NB = 1> struct TensorLds {
  static constexpr int SPAN = L::TMAXALL - L::TMINALL;
  static constexpr int WX = CX + SPAN;
  static constexpr int WXS = XQ == 1 ? WX : (WX + 1) / 2 * 2;
  static constexpr int WY = CY + SPAN;
  static constexpr int A = NTL * WY * WXS; // one window buffer
  static constexpr int X = NTL * WY * 2 * CX; // one x-pass buffer
  static constexpr int TOTAL = NB * (A + X);  // NB (1 or 2) buffers of each
};

// Combination of one point's values on ntl time levels (canonical order:
// s0*v0, then fma(s1, v1, .), fma(s2, v2, .)); with the ENO rule
// (interpolate_eno_3d_3tl.cc:146-159) a 3-level result outside the range of
// the three values falls back to the pair l0, l0+1.  nlev = 1 (a transfer
// whose target time matches level 0 exactly, run by a kernel built for NTL
// levels) gives level 0's value untouched, as a one-level transfer does.
template <int NTL>
__device__ __forceinline__ double time_combine(const double *v, const double *s, const double *s2,
                                               int l0, int teno, int nlev = NTL) {
  if constexpr (NTL == 1) {
    return v[0];
  } else if (nlev == 1) {
    return v[0];
  } else if constexpr (NTL == 2) {
    return fmad(s[1], v[1], s[0] * v[0]);
  } else {
    double d = fmad(s[2], v[2], fmad(s[1], v[1], s[0] * v[0]));
    if (teno && (d > fmax(fmax(v[0], v[1]), v[2]) || d < fmin(fmin(v[0], v[1]), v[2]))) {
      const double u = l0 ? v[1] : v[0], w = l0 ? v[2] : v[1];
      d = fmad(s2[1], w, s2[0] * u);
    }
    return d;
  }
}

template <typename TS> struct Pair;
template <> struct Pair<double> { using T = double2; };
template <> struct Pair<float> { using T = float2; };

// Precision modes: storage type S of the arrays, arithmetic type C of the
// passes.  0: FP64 (the drop-in).  1: FP32 storage, FP64 arithmetic -- the
// product of a dyadic weight (at most 9 significant bits) and an FP32 value
// is exact in FP64 and the sums carry 29 guard bits, an error-free-product
// accumulation; the time levels are combined (in FP64) before the spatial
// passes (pass_levels).  2: FP32 storage, compensated FP32 sums (Dot2 with
// FMA TwoProduct and TwoSum).  Time interpolation is FP64 in every mode.
template <int PR> struct Prec;
template <> struct Prec<0> { using S = double; using C = double; };
template <> struct Prec<1> { using S = float; using C = double; };
template <> struct Prec<2> { using S = float; using C = float; };

// line operator in the arithmetic type C (values converted from storage)
template <class L, int P, typename C, class V> __device__ __forceinline__ C line_apply(V v) {
  if constexpr (sizeof(C) == 8)
    return L::template apply<P>([&](int t) { return (double)v(t); });
  else
    return L::template apply_f<P>(v);
}

// compile-time loop: f(integral_constant<int,I>) for I = 0..N-1 while f
// returns true
template <int I, int N, class F> __device__ __forceinline__ bool unroll_while(F &f) {
  if constexpr (I < N) {
    if (!f(std::integral_constant<int, I>{}))
      return false;
    return unroll_while<I + 1, N>(f);
  }
  return true;
}

// One z-marching tile; tile b indexes (x tile, y tile, z chunk) of a.
// PR: precision mode (Prec); PF: planes loaded ahead of their use (register
// buffers per plane); XQ: coarse bases per thread in the x-pass (1, or 2 with
// paired LDS accesses); PIPE: schedule of the plane loop (0: window store,
// barrier, x-pass, barrier, y-pass per plane; 1: software pipeline -- the
// window of plane c, the x-pass of plane c-1 and the y-pass of plane c-2
// between two barriers, with double-buffered window and x-pass results);
// NSRC: source time levels; NSRC > NTL (= 1) combines the levels in FP64 as
// the window is stored (time interpolation before the spatial stencil, exact
// by linearity up to rounding; the FP32-storage, FP64-arithmetic mode), so
// the passes run once instead of once per level.
// smem must be 16-byte aligned and hold TensorLds<..., XQ, PIPE ? 2 : 1>
template <class L, int NTL, int CX, int CY, int IPT, int PR = 0, int PF = 1, int XQ = 1,
          int PIPE = 0, int NSRC = NTL>
__device__ __forceinline__ void tensor_tile(const TensorArgs &a, int b, double *smem) {
  using TS = typename Prec<PR>::S; // storage
  using CT = typename Prec<PR>::C; // arithmetic of the passes
  constexpr bool TF = NSRC > NTL;  // time levels combined at the window store
  static_assert(!TF || (NTL == 1 && sizeof(CT) == 8), "time-first needs one pass level in FP64");
  using WT = std::conditional_t<TF, CT, TS>; // LDS window type
  constexpr int FX = 2 * CX;
  constexpr int NT = FX * CY / IPT;
  constexpr int NQY = NT / FX;
  static_assert(NQY * IPT == CY, "tile rows must divide among threads");
  static_assert(XQ == 1 || (XQ == 2 && CX % 2 == 0), "x-pass bases per thread: 1 or 2");
  constexpr int NB = PIPE ? 2 : 1;
  using LD = TensorLds<L, NTL, CX, CY, XQ, NB>;
  constexpr int WX = LD::WX;
  constexpr int WXS = LD::WXS; // LDS row stride of the window
  constexpr int WY = LD::WY;
  constexpr int R = L::R;
  constexpr int NLD = (WY * WXS + NT - 1) / NT; // window elements per thread
  // planes per unrolled round: ring slot, prefetch slot and buffer parity
  // are all compile-time constants within a round
  constexpr int U0 = (R % PF == 0) ? R : R * PF;
  constexpr int U1 = (PIPE && U0 % 2) ? 2 * U0 : U0;
  constexpr int U = (U1 % PF == 0) ? U1 : U1 * PF;
  static_assert(U % PF == 0 && U % R == 0 && U % NB == 0, "unroll must cover the slots");

  // buffer B of the window (storage type) and of the x-pass (arithmetic type)
  constexpr int ABYTES = (LD::A * (int)sizeof(WT) + 15) / 16 * 16;
  constexpr int XBYTES = (LD::X * (int)sizeof(CT) + 15) / 16 * 16;
  auto Abuf = [&](int B) {
    return reinterpret_cast<WT(*)[WY][WXS]>(reinterpret_cast<char *>(smem) + B * ABYTES);
  };
  auto Xbuf = [&](int B) {
    return reinterpret_cast<CT(*)[WY][FX]>(reinterpret_cast<char *>(smem) + NB * ABYTES + B * XBYTES);
  };

  const int tx = b % a.ntiles[0];
  b /= a.ntiles[0];
  const int ty = b % a.ntiles[1];
  const int tz = b / a.ntiles[1];

  const int qx0 = a.qlo[0] + tx * CX;
  const int qy0 = a.qlo[1] + ty * CY;
  const int qz0 = a.qlo[2] + tz * a.zchunk;
  const int qz1 = min(qz0 + a.zchunk, a.qlo[2] + a.nq[2]) - 1;

  const int wx0 = qx0 + L::TMINALL, wy0 = qy0 + L::TMINALL;
  const int lx0 = max(wx0, a.flo[0]), lx1 = min(wx0 + WX - 1, a.fhi[0]);
  const int ly0 = max(wy0, a.flo[1]), ly1 = min(wy0 + WY - 1, a.fhi[1]);
  const int c0 = max(qz0 + L::TMINALL, a.flo[2]);
  const int c1 = min(qz1 + L::TMAXALL, a.fhi[2]);

  const int tid = threadIdx.x;
  const int fx = tid % FX, qyl = tid / FX;
  const int gfx = 2 * qx0 + fx - a.off[0];
  const bool xin = gfx >= 0 && gfx < a.n[0];

  // this thread's share of the plane window: offsets into the source plane
  // (or -1 outside the clipped footprint)
  long long woff[NLD];
#pragma unroll
  for (int k = 0; k < NLD; ++k) {
    const int e = tid + k * NT;
    const int r = e / WXS, x = e - r * WXS;
    const int gy = wy0 + r, gx = wx0 + x;
    woff[k] = (e < WY * WXS && x < WX && gy >= ly0 && gy <= ly1 && gx >= lx0 && gx <= lx1)
                  ? gx + a.sj * gy
                  : -1;
  }
  // windows of the next PF planes, loaded PF planes ahead of their use;
  // plane c sits in buffer (c - c0) % PF
  TS pre[PF][NSRC][NLD];
  auto fetch = [&](int c, auto slotc) {
    constexpr int SL = decltype(slotc)::value;
#pragma unroll
    for (int k = 0; k < NLD; ++k)
      if (woff[k] >= 0)
#pragma unroll
        for (int tl = 0; tl < NSRC; ++tl)
          pre[SL][tl][k] = static_cast<const TS *>(a.src[tl])[woff[k] + a.sk * c];
  };

  // ring of y-pass results: slot j holds the planes c with (c - c0) % R == j
  CT ring[NTL][R][IPT][2];
#pragma unroll
  for (int tl = 0; tl < NTL; ++tl)
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int m = 0; m < IPT; ++m)
        ring[tl][i][m][0] = ring[tl][i][m][1] = CT(0);

  auto prologue = [&](auto Sc) -> bool {
    constexpr int S = decltype(Sc)::value;
    if (c0 + S <= c1)
      fetch(c0 + S, std::integral_constant<int, S>{});
    return true;
  };
  unroll_while<0, PF>(prologue);

  // window registers pre[SL] -> LDS window buffer B
  auto store_window = [&](auto SLc, auto Bc) {
    constexpr int SL = decltype(SLc)::value;
    auto A = Abuf(decltype(Bc)::value);
#pragma unroll
    for (int k = 0; k < NLD; ++k)
      if (woff[k] >= 0) {
        const int e = tid + k * NT;
        if constexpr (TF) {
          // Lagrange time levels only (no ENO rule): canonical FP64 order
          double v[NSRC];
#pragma unroll
          for (int tl = 0; tl < NSRC; ++tl)
            v[tl] = (double)pre[SL][tl][k];
          (&A[0][0][0])[e] = time_combine<NSRC>(v, a.s, a.s2, a.l0, 0, a.nlev);
        } else {
#pragma unroll
          for (int tl = 0; tl < NTL; ++tl)
            (&A[tl][0][0])[e] = pre[SL][tl][k];
        }
      }
  };
  // x-pass: both children of each base, window buffer B -> x buffer B
  auto xpass = [&](auto Bc) {
    auto A = Abuf(decltype(Bc)::value);
    auto X = Xbuf(decltype(Bc)::value);
    // work items (time level, row, base group) spread over all threads
    if constexpr (XQ == 1) {
      for (int e = tid; e < NTL * WY * CX; e += NT) {
        const int tl = e / (WY * CX), e2 = e - tl * (WY * CX);
        const int r = e2 / CX, q = e2 - r * CX;
        const WT *w = &A[tl][r][q - L::TMINALL];
        const CT v0 = line_apply<L, 0, CT>([&](int t) { return w[L::TMIN[0] + t]; });
        const CT v1 = line_apply<L, 1, CT>([&](int t) { return w[L::TMIN[1] + t]; });
        X[tl][r][2 * q] = v0;
        X[tl][r][2 * q + 1] = v1;
      }
    } else {
      // bases q0, q0+1 from the window columns q0 .. q0+SPAN+1 (window
      // column 0 is offset TMINALL of base 0), read and written in pairs
      using P2 = typename Pair<WT>::T;
      using C2 = typename Pair<CT>::T;
      constexpr int NP = (LD::SPAN + 3) / 2;
      constexpr int NQ2 = WY * (CX / 2);
      for (int e = tid; e < NTL * NQ2; e += NT) {
        const int tl = e / NQ2, e2 = e - tl * NQ2;
        const int r = e2 / (CX / 2), q0 = 2 * (e2 - r * (CX / 2));
        WT v[2 * NP];
#pragma unroll
        for (int m = 0; m < NP; ++m) {
          const P2 p = *reinterpret_cast<const P2 *>(&A[tl][r][q0 + 2 * m]);
          v[2 * m] = p.x;
          v[2 * m + 1] = p.y;
        }
        constexpr int O0 = L::TMIN[0] - L::TMINALL, O1 = L::TMIN[1] - L::TMINALL;
        C2 lo, hi;
        lo.x = line_apply<L, 0, CT>([&](int t) { return v[O0 + t]; });
        lo.y = line_apply<L, 1, CT>([&](int t) { return v[O1 + t]; });
        hi.x = line_apply<L, 0, CT>([&](int t) { return v[1 + O0 + t]; });
        hi.y = line_apply<L, 1, CT>([&](int t) { return v[1 + O1 + t]; });
        *reinterpret_cast<C2 *>(&X[tl][r][2 * q0]) = lo;
        *reinterpret_cast<C2 *>(&X[tl][r][2 * q0 + 2]) = hi;
      }
    }
  };
  // y-pass of plane p (x buffer B) into ring slot J, then the z-emission of
  // the fine planes whose last coarse plane is p: their planes p-NTP+1..p
  // sit in slots (J-NTP+1+t) mod R
  auto ypass_emit = [&](auto Bc, auto Jc, int p) {
    constexpr int J = decltype(Jc)::value;
    auto X = Xbuf(decltype(Bc)::value);
#pragma unroll
    for (int m = 0; m < IPT; ++m) {
      const int qy = qyl + NQY * m;
#pragma unroll
      for (int tl = 0; tl < NTL; ++tl) {
        const CT *col = &X[tl][qy - L::TMINALL][fx];
        ring[tl][J][m][0] = line_apply<L, 0, CT>([&](int t) { return col[(L::TMIN[0] + t) * FX]; });
        ring[tl][J][m][1] = line_apply<L, 1, CT>([&](int t) { return col[(L::TMIN[1] + t) * FX]; });
      }
    }
    auto emit = [&](auto PZc) {
      constexpr int PZ = decltype(PZc)::value;
      constexpr int NTP = L::NTAP[PZ];
      constexpr int TMAX = L::TMIN[PZ] + NTP - 1;
      const int base = p - TMAX;
      if (base < qz0 || base > qz1)
        return;
      const int fz = 2 * base + PZ - a.off[2];
      if (fz < 0 || fz >= a.n[2])
        return;
#pragma unroll
      for (int m = 0; m < IPT; ++m) {
#pragma unroll
        for (int py = 0; py < 2; ++py) {
          const int fy = 2 * (qy0 + qyl + NQY * m) + py - a.off[1];
          double v[NTL];
#pragma unroll
          for (int tl = 0; tl < NTL; ++tl)
            v[tl] = (double)line_apply<L, PZ, CT>(
                [&](int t) { return ring[tl][(J - NTP + 1 + t + 2 * R) % R][m][py]; });
          // time interpolation in FP64 in every precision mode
          const double out = time_combine<NTL>(v, a.s, a.s2, a.l0, a.teno, a.nlev);
          if (xin && fy >= 0 && fy < a.n[1])
            __builtin_nontemporal_store((TS)out, &static_cast<TS *>(a.dst)[gfx + a.dj * fy + a.dk * fz]);
        }
      }
    };
    emit(std::integral_constant<int, 0>{});
    emit(std::integral_constant<int, 1>{});
  };

  int c = c0;
  if constexpr (PIPE == 0) {
    auto plane = [&](auto Uc) -> bool {
      constexpr int J = decltype(Uc)::value % R;  // ring slot of plane c
      constexpr int SL = decltype(Uc)::value % PF; // prefetch buffer of plane c
      if (c > c1)
        return false;
      store_window(std::integral_constant<int, SL>{}, std::integral_constant<int, 0>{});
      if (c + PF <= c1)
        fetch(c + PF, std::integral_constant<int, SL>{});
      __syncthreads();
      xpass(std::integral_constant<int, 0>{});
      __syncthreads();
      ypass_emit(std::integral_constant<int, 0>{}, std::integral_constant<int, J>{}, c);
      ++c;
      return true;
    };
    while (unroll_while<0, U>(plane)) {
    }
  } else {
    // iteration c: window of plane c (buffer c&1), x-pass of plane c-1
    // (buffer (c-1)&1), y-pass of plane c-2 (buffer c&1), one barrier.
    // Buffer B is written by the window store of iteration c and read by
    // the x-pass of iteration c+1; its x buffer is written there and read
    // by the y-pass of iteration c+2: every hazard crosses a barrier.
    auto iter = [&](auto Uc) -> bool {
      constexpr int u = decltype(Uc)::value;
      constexpr int B0 = u & 1, B1 = (u + 1) & 1;
      constexpr int J2 = ((u - 2) % R + R) % R; // ring slot of plane c-2
      constexpr int SL = u % PF;                 // prefetch buffer of plane c
      if (c > c1 + 2)
        return false;
      if (c <= c1) {
        store_window(std::integral_constant<int, SL>{}, std::integral_constant<int, B0>{});
        if (c + PF <= c1)
          fetch(c + PF, std::integral_constant<int, SL>{});
      }
      if (c - 1 >= c0 && c - 1 <= c1)
        xpass(std::integral_constant<int, B1>{});
      if (c - 2 >= c0)
        ypass_emit(std::integral_constant<int, B0>{}, std::integral_constant<int, J2>{}, c - 2);
      __syncthreads();
      ++c;
      return true;
    };
    while (unroll_while<0, U>(iter)) {
    }
  }
  // the next tile of a persistent workgroup overwrites the LDS window
  __syncthreads();
}

// ---- pointwise region operators -------------------------------------------
__device__ __forceinline__ double minmod(double a, double b) {
  if (a * b < 0)
    return 0.0;
  return __builtin_fabs(a) < __builtin_fabs(b) ? a : b;
}

// prolongate_3d_o5_monotone_rf2 at fine point (i,j,k) from one source
// level (prolongate_3d_o5_monotone_rf2.cc:45-791): the order-5 stencil
// (3,-25,150,150,-25,3)/256 on every axis where the point lies between
// coarse points, summed flat in the reference's order -- one axis: its six
// taps; two axes: the outer axis' offsets outside, the inner axis' inside;
// three axes: one flat sum per z offset, the six added in z order -- each
// product and sum rounded on its own; where the value leaves the range of
// the 2, 4 or 8 nearest coarse values, their average (weights 1/2 per axis,
// the same corner order) replaces it
__device__ __forceinline__ double mono5_value(const double *src, long long J, long long K, int gi, int gj,
                                              int gk) {
  constexpr double f[6] = {3 / 256.0, -25 / 256.0, 150 / 256.0, 150 / 256.0, -25 / 256.0, 3 / 256.0};
  const double *p = src + (gi >> 1) + J * (gj >> 1) + K * (gk >> 1);
  const int pi = gi & 1, pj = gj & 1, pk = gk & 1;
  const int nodd = pi + pj + pk;
  if (nodd == 0)
    return p[0];
  // odd axes in x, y, z order: strides s[0] (inner) .. s[nodd-1] (outer)
  long long s[3];
  int m = 0;
  if (pi)
    s[m++] = 1;
  if (pj)
    s[m++] = J;
  if (pk)
    s[m++] = K;
  if (nodd == 1) {
    const long long a = s[0];
    double r = f[0] * p[-2 * a];
#pragma unroll
    for (int t = 1; t < 6; ++t)
      r = r + f[t] * p[(t - 2) * a];
    const double c0 = p[0], c1 = p[a];
    if (r > fmax(c0, c1) || r < fmin(c0, c1))
      r = 0.5 * c0 + 0.5 * c1;
    return r;
  }
  if (nodd == 2) {
    const long long a = s[0], b = s[1];
    double r = (f[0] * f[0]) * p[-2 * a - 2 * b];
    for (int tb = 0; tb < 6; ++tb)
#pragma unroll
      for (int ta = 0; ta < 6; ++ta)
        if (ta | tb)
          r = r + (f[ta] * f[tb]) * p[(ta - 2) * a + (tb - 2) * b];
    const double c00 = p[0], c10 = p[a], c01 = p[b], c11 = p[a + b];
    if (r > fmax(fmax(c00, c10), fmax(c01, c11)) || r < fmin(fmin(c00, c10), fmin(c01, c11))) {
      const double w = 0.5 * 0.5;
      r = w * c00 + w * c10;
      r = r + w * c01;
      r = r + w * c11;
    }
    return r;
  }
  const long long a = s[0], b = s[1], c = s[2];
  double r = 0.0;
  for (int tc = 0; tc < 6; ++tc) {
    double res = (f[0] * f[0] * f[tc]) * p[-2 * a - 2 * b + (tc - 2) * c];
    for (int tb = 0; tb < 6; ++tb)
#pragma unroll
      for (int ta = 0; ta < 6; ++ta)
        if (ta | tb)
          res = res + (f[ta] * f[tb] * f[tc]) * p[(ta - 2) * a + (tb - 2) * b + (tc - 2) * c];
    r = tc == 0 ? res : r + res;
  }
  const double c000 = p[0], c100 = p[a], c010 = p[b], c001 = p[c], c110 = p[a + b], c101 = p[a + c],
               c011 = p[b + c], c111 = p[a + b + c];
  const double mx = fmax(fmax(fmax(c000, c100), fmax(c010, c001)), fmax(fmax(c110, c101), fmax(c011, c111)));
  const double mn = fmin(fmin(fmin(c000, c100), fmin(c010, c001)), fmin(fmin(c110, c101), fmin(c011, c111)));
  if (r > mx || r < mn) {
    const double w = 0.5 * 0.5 * 0.5;
    double r1 = w * c000 + w * c100;
    r1 = r1 + w * c010;
    r1 = r1 + w * c110;
    double r2 = w * c001 + w * c101;
    r2 = r2 + w * c011;
    r2 = r2 + w * c111;
    r = r1 + r2;
  }
  return r;
}

// Staggered operators (prolongate_3d_stagger{011,101,110,111}.cc,
// restrict_3d_stagger{011,101,110,111}.cc).  Bit d of stag is set where the
// field is staggered in direction d.  Per direction and fine parity f the
// prolongation uses one tap set (coeffs_prolongate_3d_stagger.hh), with the
// first tap at coarse base + stag_lo and stag_n taps:
//   kind 0, staggered, f = 0: coeff[1] (to b - 1/4)
//   kind 1, staggered, f = 1: coeff[2] (to b + 1/4)
//   kind 2, unstaggered, f = 1: coeff[3] (to b + 1/2)
//   kind 3, unstaggered, f = 0: coeff[4] for order 3; order 2 takes the
//        coarse value (the reference's two-dimensional sums)
__device__ __forceinline__ int stag_kind(int stag, int d, int f) {
  return ((stag >> d) & 1) ? f : (f ? 2 : 3);
}
__device__ __forceinline__ int stag_lo(int order, int kind) {
  return order == 3 && kind == 0 ? -2 : order == 2 && kind == 3 ? 0 : -1;
}
__device__ __forceinline__ int stag_n(int order, int kind) {
  return order == 2 ? (kind == 3 ? 1 : 3) : (kind == 3 ? 3 : 4);
}
__device__ __forceinline__ double stag_coef(int order, int kind, int t) {
  if (order == 2) {
    switch (kind) {
    case 0: return t == 0 ? 5.0 / 32.0 : t == 1 ? 15.0 / 16.0 : -3.0 / 32.0;
    case 1: return t == 0 ? -3.0 / 32.0 : t == 1 ? 15.0 / 16.0 : 5.0 / 32.0;
    case 2: return t == 0 ? -1.0 / 8.0 : t == 1 ? 3.0 / 4.0 : 3.0 / 8.0;
    }
    return 1.0;
  }
  switch (kind) {
  case 0: return t == 0 ? -5.0 / 128.0 : t == 1 ? 35.0 / 128.0 : t == 2 ? 105.0 / 128.0 : -7.0 / 128.0;
  case 1: return t == 0 ? -7.0 / 128.0 : t == 1 ? 105.0 / 128.0 : t == 2 ? 35.0 / 128.0 : -5.0 / 128.0;
  case 2: return t == 0 ? -5.0 / 64.0 : t == 1 ? 37.0 / 64.0 : t == 2 ? 37.0 / 64.0 : -5.0 / 64.0;
  }
  return t == 1 ? 34.0 / 32.0 : -1.0 / 32.0;
}

// one staggered prolongation value at fine point (gi,gj,gk) relative to the
// source array: sum from 0 of coeff_x * coeff_y * coeff_z * src, x taps
// outermost and z innermost, each operation rounded as written (the
// coefficient products are exact: dyadic numerators below 2^21)
__device__ __forceinline__ double stagger_value(const double *src, long long J, long long K, int gi, int gj,
                                                int gk, int stag, int order) {
  const int kx = stag_kind(stag, 0, gi & 1), ky = stag_kind(stag, 1, gj & 1), kz = stag_kind(stag, 2, gk & 1);
  const int nx = stag_n(order, kx), ny = stag_n(order, ky), nz = stag_n(order, kz);
  const double *p = src + ((gi >> 1) + stag_lo(order, kx)) + J * ((gj >> 1) + stag_lo(order, ky)) +
                    K * ((gk >> 1) + stag_lo(order, kz));
  double r = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if (a >= nx)
      break;
    const double cx = stag_coef(order, kx, a);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (b >= ny)
        break;
      const double cxy = cx * stag_coef(order, ky, b);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c >= nz)
          break;
        const double t = (cxy * stag_coef(order, kz, c)) * p[a + J * b + K * c];
        r = r + t;
      }
    }
  }
  return r;
}

// one staggered restriction value; p is the fine point 2i,2j,2k (the
// coarse point's coincident child): staggered directions take fine points
// 2i-1..2i+2 with {-1, 9, 9, -1}/16, unstaggered ones 2i-1..2i+1 with
// {1, 6, 1}/8 (coeffs_restrict_3d_stagger.hh, ORDER_STAG 3)
__device__ __forceinline__ double stagger_restrict_value(const double *p, long long J, long long K, int stag) {
  auto coef = [](int s, int t) -> double {
    if (s)
      return t == 0 || t == 3 ? -1.0 / 16.0 : 9.0 / 16.0;
    return t == 1 ? 6.0 / 8.0 : 1.0 / 8.0;
  };
  const int sx = stag & 1, sy = (stag >> 1) & 1, sz = (stag >> 2) & 1;
  const double *q = p - 1 - J - K;
  double r = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    if (a >= 3 + sx)
      break;
    const double cx = coef(sx, a);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (b >= 3 + sy)
        break;
      const double cxy = cx * coef(sy, b);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c >= 3 + sz)
          break;
        const double t = (cxy * coef(sz, c)) * q[a + J * b + K * c];
        r = r + t;
      }
    }
  }
  return r;
}

// value of destination point (i,j,k) of a pointwise operator (loads only,
// so that a caller can issue the loads of several points before storing)
template <int OP>
__device__ __forceinline__ double point_value(const PointArgs &a, int i, int j, int k) {
  static_assert(OP != PO_TVD, "TVD runs per coarse cell (tvd_cell)");
  double out;
  if constexpr (OP == PO_COPY) {
    out = a.src[0][i + a.sj * j + a.sk * k];
  } else if constexpr (OP == PO_RESTRICT_VC) {
    out = a.src[0][2 * i + a.sj * (2 * j) + a.sk * (2 * k)];
  } else if constexpr (OP == PO_RESTRICT_CC) {
    // restrict_3d_cc_rf2.cc:118-126: eight children, fixed order, weight 1/8
    const double *p = a.src[0] + 2 * i + a.sj * (2 * j) + a.sk * (2 * k);
    const long long J = a.sj, K = a.sk;
    const double w = (1.0 / 2) * (1.0 / 2) * (1.0 / 2);
    double r = w * p[0];
    r = r + w * p[1];
    r = r + w * p[J];
    r = r + w * p[1 + J];
    r = r + w * p[K];
    r = r + w * p[1 + K];
    r = r + w * p[J + K];
    r = r + w * p[1 + J + K];
    out = r;
  } else if constexpr (OP == PO_MONO5) {
    // each time level prolongated, then combined in time (data.cc:497-521)
    const int gi = a.off[0] + i, gj = a.off[1] + j, gk = a.off[2] + k;
    double v[3];
    v[0] = mono5_value(a.src[0], a.sj, a.sk, gi, gj, gk);
    if (a.ntl == 1) {
      out = v[0];
    } else {
      v[1] = mono5_value(a.src[1], a.sj, a.sk, gi, gj, gk);
      if (a.ntl == 2) {
        out = time_combine<2>(v, a.s, a.s2, a.l0, 0);
      } else {
        v[2] = mono5_value(a.src[2], a.sj, a.sk, gi, gj, gk);
        out = time_combine<3>(v, a.s, a.s2, a.l0, a.teno);
      }
    }
  } else if constexpr (OP == PO_STAGGER) {
    // each time level prolongated, then combined in time (data.cc:497-521)
    const int gi = a.off[0] + i, gj = a.off[1] + j, gk = a.off[2] + k;
    double v[3];
    v[0] = stagger_value(a.src[0], a.sj, a.sk, gi, gj, gk, a.stag, a.sorder);
    if (a.ntl == 1) {
      out = v[0];
    } else {
      v[1] = stagger_value(a.src[1], a.sj, a.sk, gi, gj, gk, a.stag, a.sorder);
      if (a.ntl == 2) {
        out = time_combine<2>(v, a.s, a.s2, a.l0, 0);
      } else {
        v[2] = stagger_value(a.src[2], a.sj, a.sk, gi, gj, gk, a.stag, a.sorder);
        out = time_combine<3>(v, a.s, a.s2, a.l0, 0);
      }
    }
  } else if constexpr (OP == PO_RESTRICT_VC2) {
    // restrict_3d_vc_rf2.cc:26-97: nested half-sums, x inside y inside z; a
    // cell-centred direction (bit of a.stag) averages its two children, a
    // vertex-centred one takes the first; the halves are exact, so each
    // half*a + half*b is one rounding however it is contracted
    const double *p = a.src[0] + 2 * i + a.sj * (2 * j) + a.sk * (2 * k);
    const int cm = a.stag;
    auto r1 = [&](const double *q) { return (cm & 1) ? 0.5 * q[0] + 0.5 * q[1] : q[0]; };
    auto r2 = [&](const double *q) { return (cm & 2) ? 0.5 * r1(q) + 0.5 * r1(q + a.sj) : r1(q); };
    out = (cm & 4) ? 0.5 * r2(p) + 0.5 * r2(p + a.sk) : r2(p);
  } else if constexpr (OP == PO_RESTRICT_STAGGER) {
    out = stagger_restrict_value(a.src[0] + 2 * i + a.sj * (2 * j) + a.sk * (2 * k), a.sj, a.sk, a.stag);
  } else if constexpr (OP == PO_RESTRICT_CC_O3) {
    // restrict_3d_cc_o3_rf2.cc:146-211: 64 terms, (w_a w_b w_c)/4096 with
    // w = {-1, 9, 9, -1}, a fastest, then b, then c; fine cells 2i-1..2i+2
    const double *p = a.src[0] + 2 * i + a.sj * (2 * j) + a.sk * (2 * k);
    constexpr double w[4] = {-1, 9, 9, -1};
    double r = 0.0;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int aa = 0; aa < 4; ++aa) {
          const double cf = w[aa] * w[b] * w[c] / 4096;
          const double x = p[(aa - 1) + a.sj * (b - 1) + a.sk * (c - 1)];
          r = (aa | b | c) == 0 ? cf * x : fmad(cf, x, r);
        }
    out = r;
  } else if constexpr (OP == PO_RESTRICT_CC_O5) {
    // restrict_3d_cc_o5_rf2.cc:36-74: 6-point sums, x inside y inside z,
    // weights {3, -25, 150, 150, -25, 3}/256; fine cells 2i-2..2i+3
    const double *p = a.src[0] + 2 * i + a.sj * (2 * j) + a.sk * (2 * k);
    constexpr double w[6] = {3.0 / 256, -25.0 / 256, 150.0 / 256, 150.0 / 256, -25.0 / 256, 3.0 / 256};
    double r2 = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double r1 = 0.0;
#pragma unroll
      for (int b = 0; b < 6; ++b) {
        const double *q = p + a.sj * (b - 2) + a.sk * (c - 2);
        double x = w[0] * q[-2];
#pragma unroll
        for (int aa = 1; aa < 6; ++aa)
          x = fmad(w[aa], q[aa - 2], x);
        r1 = b == 0 ? w[0] * x : fmad(w[b], x, r1);
      }
      r2 = c == 0 ? w[0] * r1 : fmad(w[c], r1, r2);
    }
    out = r2;
  } else if constexpr (OP == PO_INTERP) {
    const long long o = i + a.sj * j + a.sk * k;
    if (a.ntl > 3) {
      // interpolate_3d_4tl / _5tl: s1*a, then fma(s_l, level l, .)
      double r = a.s[0] * a.src[0][o];
      r = fmad(a.s[1], a.src[1][o], r);
      r = fmad(a.s[2], a.src[2][o], r);
      r = fmad(a.s[3], a.src[3][o], r);
      if (a.ntl == 5)
        r = fmad(a.s[4], a.src[4][o], r);
      out = r;
    } else if (a.ntl == 3) {
      const double v[3] = {a.src[0][o], a.src[1][o], a.src[2][o]};
      out = time_combine<3>(v, a.s, a.s2, a.l0, a.teno);
    } else if (a.ntl == 1) {
      // a batched same-level interpolation whose target time has become an
      // exact match: the copy of that level
      out = a.src[0][o];
    } else {
      const double v[2] = {a.src[0][o], a.src[1][o]};
      out = time_combine<2>(v, a.s, a.s2, a.l0, a.teno);
    }
  }
  return out;
}

template <int OP>
__device__ __forceinline__ void point_elem(const PointArgs &a, int i, int j, int k) {
  a.dst[i + a.dj * j + a.dk * k] = point_value<OP>(a, i, j, k);
}

// Cell-centred TVD prolongation (prolongate_3d_cc_real8_tvd.F90:183-225)
// of all children of one coarse cell.  Fine point f of axis d has parent
// (off+f)>>1 and side (off+f)&1, so coarse cell q owns the children
// f = 2q - off + s, s = 0, 1, those inside [0, n).  Per axis the slope is
// minmod(c - c[-1], c[+1] - c), signed by the side and scaled by 1/4 (both
// exact); every child is ((c + sx) + sy) + sz, the reference's order.  The
// seven loads serve up to eight children.
__device__ __forceinline__ void tvd_range(const PointArgs &a, int d, int *q0, int *nq) {
  *q0 = a.off[d] >> 1;
  *nq = ((a.off[d] + a.n[d] - 1) >> 1) - *q0 + 1;
}

// parent value and the three minmod slopes of coarse cell (qx,qy,qz) on
// time level tl
typedef double dbl2 __attribute__((ext_vector_type(2)));

struct TvdCell {
  double c, mx, my, mz;
};
__device__ __forceinline__ TvdCell tvd_load(const PointArgs &a, int tl, int qx, int qy, int qz) {
  const long long J = a.sj, K = a.sk;
  const double *p = a.src[tl] + qx + J * qy + K * qz;
  const double c = p[0];
  return TvdCell{c, minmod(c - p[-1], p[1] - c), minmod(c - p[-J], p[J] - c),
                 minmod(c - p[-K], p[K] - c)};
}

// the children of one cell on NTL time levels (data<T>::transfer_time: each
// level prolongated, then the levels combined in time)
template <int NTL>
__device__ __forceinline__ void tvd_store(const PointArgs &a, int qx, int qy, int qz,
                                          const TvdCell *v) {
  const double q = 1.0 / 4;
  const int fx0 = 2 * qx - a.off[0], fy0 = 2 * qy - a.off[1], fz0 = 2 * qz - a.off[2];
#pragma unroll
  for (int z = 0; z < 2; ++z) {
    const int fz = fz0 + z;
    if (fz < 0 || fz >= a.n[2])
      continue;
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int fy = fy0 + y;
      if (fy < 0 || fy >= a.n[1])
        continue;
      double out[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) {
        double w[NTL];
#pragma unroll
        for (int tl = 0; tl < NTL; ++tl) {
          const double sx = (x ? q : -q) * v[tl].mx, sy = (y ? q : -q) * v[tl].my,
                       sz = (z ? q : -q) * v[tl].mz;
          w[tl] = v[tl].c + sx + sy + sz;
        }
        out[x] = time_combine<NTL>(w, a.s, a.s2, a.l0, a.teno);
      }
      double *p = &a.dst[fx0 + a.dj * fy + a.dk * fz];
      const bool in0 = fx0 >= 0 && fx0 < a.n[0], in1 = fx0 + 1 >= 0 && fx0 + 1 < a.n[0];
      // both x children in one 16-byte store where the pair is aligned, so a
      // wavefront writes whole contiguous lines (two 8-byte stores per lane
      // at a 16-byte stride leave every line half written per instruction)
      if (in0 && in1 && ((reinterpret_cast<uintptr_t>(p) & 15) == 0)) {
        __builtin_nontemporal_store(dbl2{out[0], out[1]}, reinterpret_cast<dbl2 *>(p));
      } else {
        if (in0)
          __builtin_nontemporal_store(out[0], p);
        if (in1)
          __builtin_nontemporal_store(out[1], p + 1);
      }
    }
  }
}

template <int NTL>
__device__ __forceinline__ void tvd_cell(const PointArgs &a, int qx, int qy, int qz) {
  TvdCell v[NTL];
#pragma unroll
  for (int tl = 0; tl < NTL; ++tl)
    v[tl] = tvd_load(a, tl, qx, qy, qz);
  tvd_store<NTL>(a, qx, qy, qz, v);
}

// tile shapes: one z-marching column of 2CX x 2CY fine points, 2CX*CY/IPT
// threads.  With 2-3 time levels the register ring is 2-3x larger, so each
// thread keeps one base row (IPT 1, 512 threads); 8 base rows per tile cut
// the y halo of the x-pass and of the window loads to (8+span)/8.
// XQ = 2: the x-pass reads and writes LDS in 16-byte pairs (two bases per
// thread), a third fewer LDS reads than one base per thread.
// SH = 0 ("wide", 32 x 8 bases) suits large regions; SH = 1 ("narrow",
// 16 x 16 bases) wastes fewer lanes on regions of a few dozen points per
// axis and has a smaller x halo per fine point.  PR: precision mode; the
// FP32-storage, FP64-arithmetic mode (one pass level) runs its wide tiles
// 64 bases across, so a fine row is 512 bytes of FP32 and the x halo is
// half as large per point.
template <int NTL, int SH = 0, int PR = 0> struct TileShape : TileShape<NTL, SH, 0> {};
template <> struct TileShape<1, 0, 0> {
  static constexpr int CX = 32, CY = 8, IPT = 2, XQ = 2;
};
template <> struct TileShape<2, 0, 0> {
  static constexpr int CX = 32, CY = 8, IPT = 1, XQ = 2;
};
template <> struct TileShape<3, 0, 0> {
  static constexpr int CX = 32, CY = 8, IPT = 1, XQ = 2;
};
template <> struct TileShape<1, 1, 0> {
  static constexpr int CX = 16, CY = 16, IPT = 2, XQ = 2;
};
template <> struct TileShape<2, 1, 0> {
  static constexpr int CX = 16, CY = 16, IPT = 1, XQ = 2;
};
template <> struct TileShape<3, 1, 0> {
  static constexpr int CX = 16, CY = 16, IPT = 1, XQ = 2;
};
template <> struct TileShape<1, 0, 1> {
  static constexpr int CX = 64, CY = 8, IPT = 2, XQ = 2;
};
// small tiles (FP64 batch only) for regions a few fine points across in x
// and in y or z: edges and corners of ghost and buffer shells, faces thin
// in x, where a 32x8 or 16x16 tile keeps a tenth of its lanes busy or less
template <int NTL> struct TileShape<NTL, 2, 0> {
  static constexpr int CX = 8, CY = 8, IPT = 1, XQ = 2;
};
// slim tiles (FP64 batch only), 4 x 16 bases, for regions a few fine points
// across in x and long in y: ghost and buffer faces normal to x, where an
// 8x8 tile leaves most of its x lanes and window columns idle
template <int NTL> struct TileShape<NTL, 3, 0> {
  static constexpr int CX = 4, CY = 16, IPT = 1, XQ = 2;
};
// the pipelined schedule (double-buffered window and x-pass results) where
// its LDS fits the 64 KB of a static allocation, else two barriers per plane
template <class L, int NTL, int CX, int CY, int XQ> constexpr int pipe_fits() {
  return TensorLds<L, NTL, CX, CY, XQ, 2>::TOTAL * 8 <= 65536 ? 1 : 0;
}
template <class L, int NTL, int CX, int CY, int XQ> constexpr int tile_lds() {
  return TensorLds<L, NTL, CX, CY, XQ, pipe_fits<L, NTL, CX, CY, XQ>() ? 2 : 1>::TOTAL;
}

// time levels the tile passes carry: the FP32-storage, FP64-arithmetic mode
// (1) combines the source levels as the window is stored
constexpr int pass_levels(int ntl, int prec) { return prec == 1 ? 1 : ntl; }

template <int NTL, int SH = 0, int PR = 0> constexpr int tile_threads() {
  using T = TileShape<NTL, SH, PR>;
  return 2 * T::CX * T::CY / T::IPT;
}
// host view of the tile shape of a transfer class: ntl source levels,
// precision mode prec, shape sh
inline void tile_shape(int ntl, int prec, int *cx, int *cy, int *threads, int sh = 0) {
  auto get = [&](auto ntlc, auto shc, auto prc) {
    constexpr int N = decltype(ntlc)::value, S = decltype(shc)::value, P = decltype(prc)::value;
    *cx = TileShape<N, S, P>::CX;
    *cy = TileShape<N, S, P>::CY;
    *threads = tile_threads<N, S, P>();
  };
  auto by_sh = [&](auto ntlc, auto prc) {
    if (sh == 0)
      get(ntlc, std::integral_constant<int, 0>{}, prc);
    else if (sh == 1)
      get(ntlc, std::integral_constant<int, 1>{}, prc);
    else if (sh == 2)
      get(ntlc, std::integral_constant<int, 2>{}, prc);
    else
      get(ntlc, std::integral_constant<int, 3>{}, prc);
  };
  auto by_prec = [&](auto ntlc) {
    if (prec == 1)
      by_sh(ntlc, std::integral_constant<int, 1>{});
    else if (prec == 2)
      by_sh(ntlc, std::integral_constant<int, 2>{});
    else
      by_sh(ntlc, std::integral_constant<int, 0>{});
  };
  const int np = pass_levels(ntl, prec);
  if (np == 1)
    by_prec(std::integral_constant<int, 1>{});
  else if (np == 2)
    by_prec(std::integral_constant<int, 2>{});
  else
    by_prec(std::integral_constant<int, 3>{});
}

} // namespace cg

namespace cg {
// Calls fn(Line<KIND,O>{}) for a runtime order; returns -1 if there is no
// such line operator.
template <int KIND, class Fn> int with_order(int order, Fn &&fn) {
  if constexpr (KIND == LK_VLAG) {
    switch (order) {
    case 1: return fn(Line<LK_VLAG, 1>{});
    case 3: return fn(Line<LK_VLAG, 3>{});
    case 5: return fn(Line<LK_VLAG, 5>{});
    case 7: return fn(Line<LK_VLAG, 7>{});
    case 9: return fn(Line<LK_VLAG, 9>{});
    case 11: return fn(Line<LK_VLAG, 11>{});
    }
  } else if constexpr (KIND == LK_CLAG) {
    switch (order) {
    case 0: return fn(Line<LK_CLAG, 0>{});
    case 1: return fn(Line<LK_CLAG, 1>{});
    case 2: return fn(Line<LK_CLAG, 2>{});
    case 3: return fn(Line<LK_CLAG, 3>{});
    case 4: return fn(Line<LK_CLAG, 4>{});
    case 5: return fn(Line<LK_CLAG, 5>{});
    }
  } else if constexpr (KIND == LK_ENO) {
    switch (order) {
    case 2: return fn(Line<LK_ENO, 2>{});
    case 3: return fn(Line<LK_ENO, 3>{});
    }
  } else if constexpr (KIND == LK_VENO) {
    if (order == 3)
      return fn(Line<LK_VENO, 3>{});
  }
  return -1;
}

// Calls fn(Line<KIND,O>{}) for a runtime (kind, order); returns -1 if there
// is no such line operator.
template <class Fn> int with_line(int kind, int order, Fn &&fn) {
  switch (kind) {
  case LK_VLAG: return with_order<LK_VLAG>(order, fn);
  case LK_CLAG: return with_order<LK_CLAG>(order, fn);
  case LK_ENO: return with_order<LK_ENO>(order, fn);
  case LK_VENO: return with_order<LK_VENO>(order, fn);
  }
  return -1;
}
} // namespace cg
