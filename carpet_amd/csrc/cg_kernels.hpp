// cg_kernels.hpp -- kernel argument blocks and launcher declarations shared
// by the per-call entry points, the transfer dispatch and the batch path.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace cg {

// Tensor-product prolongation (Lagrange vertex/cell, ENO cell) fused with
// time interpolation over up to 3 source time levels.
struct TensorArgs {
  const void *src[3]; // per time level: element (0,0,0) of the source array
  long long sj, sk;   // source strides (elements)
  void *dst;          // destination element at the region origin
  long long dj, dk;   // destination strides (elements)
  int n[3];           // region extents (fine points)
  int off[3];         // fine index of the region origin relative to the source
                      // array: fine point f has coarse base (off+f)>>1 and
                      // parity (off+f)&1
  int flo[3], fhi[3]; // exact coarse footprint of the region (inclusive)
  int qlo[3];         // first coarse base (off>>1)
  int nq[3];          // number of coarse bases per axis
  int ntiles[3];      // tile counts: x, y (in CX/CY bases), z (chunks)
  int zchunk;         // coarse bases per z chunk
  double s[3];        // time weights
  double s2[2];       // ENO time rule: weights of the fallback pair
  int l0;             //   first level of the fallback pair (0: 1-2, 1: 2-3)
  int teno;           //   1: interpolate_eno_3d_3tl's rule for 3 levels
  int nlev;           // time levels this transfer combines: the kernel's NTL,
                      // or 1 (an exact time match inside a class of NTL
                      // levels: the value of level 0, every src[] = src[0])
};

// Division by a fixed d >= 1 for dividends n < 2^31 as a multiply and a
// shift: m = ceil(2^(31+l) / d), l = ceil(log2 d), q = (n * m) >> (31 + l)
// (Granlund & Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, Thm 4.2 with N = 31: m*d - 2^(31+l) < 2^l).
// m < 2^32 for every d, so the product is one 32x32->64 multiply.
struct FastDiv {
  unsigned m;
  int s;
};
inline FastDiv make_fastdiv(unsigned d) {
  int l = 0;
  while ((1ull << l) < d)
    ++l;
  return FastDiv{(unsigned)(((1ull << (31 + l)) + d - 1) / d), 31 + l};
}
__device__ __forceinline__ unsigned fastdiv(unsigned n, FastDiv f) {
  return (unsigned)(((unsigned long long)n * f.m) >> f.s);
}

// Pointwise region operators: copy (with slab shift), injection
// restriction, cell-centred restriction, time interpolation, cell TVD.
enum PointOp { PO_COPY = 0, PO_RESTRICT_VC = 1, PO_RESTRICT_CC = 2,
               PO_INTERP = 3, PO_TVD = 4, PO_RESTRICT_CC_O3 = 5, PO_RESTRICT_CC_O5 = 6,
               PO_MONO5 = 7, PO_STAGGER = 8, PO_RESTRICT_STAGGER = 9, PO_RESTRICT_VC2 = 10 };

struct PointArgs {
  const double *src[5]; // source element matching region point (0,0,0)
                        // (copy/interp); for restrict: first child; for
                        // TVD: source element (0,0,0) of the array
  long long sj, sk;
  double *dst;          // destination element at the region origin
  long long dj, dk;
  int n[3];             // region extents (destination points)
  int ntl;              // time levels (interp: 2..5; TVD and the other
                        // prolongations: 1..3)
  double s[5];          // time weights
  double s2[2];         // ENO time rule: fallback pair weights
  int l0, teno;         //   first level of the pair; 1 = rule on (3 levels)
  int off[3];           // TVD: fine offsets (off+f)/2 = parent, (off+f)%2 = side
  int stag;             // staggered operators: bit d set = staggered in direction d;
                        // restrict_3d_vc_rf2: bit d set = direction d cell-centred
  int sorder;           //   prolongation order (2 or 3)
  FastDiv fx, fxy;      // division by n[0] and n[0]*n[1] (batch work items)
};

// tile configuration of the tensor kernel
struct TileCfg {
  int cx, cy; // coarse bases per tile in x and y (fine tile 2cx x 2cy)
  int zchunk; // coarse bases per z chunk
};

// ---- launchers (defined in cg_kernels.hip) --------------------------------
// kind: LK_VLAG/LK_CLAG/LK_ENO; prec 0: double sources/destination,
// 1: float sources/destination with compensated accumulation.
// per line kind (LK_*); defined in cg_tensor_plain.hpp, instantiated in
// cg_tensor_{vlag,clag,eno}.hip
template <int KIND>
int launch_tensor_kind(int order, int ntl, int prec, const TensorArgs &a, const TileCfg &cfg,
                       hipStream_t s);
int launch_tensor(int kind, int order, int ntl, int prec, const TensorArgs &a,
                  const TileCfg &cfg, hipStream_t s);
// computes ntiles/zchunk of a for a tile configuration; returns block count
long long tensor_tiles(TensorArgs &a, const TileCfg &cfg);
int launch_point(int op, const PointArgs &a, hipStream_t s);
int launch_fill(double *dst, double v, size_t n, hipStream_t s);
int launch_stream_copy(double *dst, const double *src, size_t n, hipStream_t s);
int launch_fill_fourier(double *d, long long sj, long long sk, const int *ext,
                        const int *lower, const int *stride, double scale, double amp,
                        int nmodes, const int *kvec, const double *phase, hipStream_t s);
int launch_fill_hash(double *d, long long sj, long long sk, const int *ext,
                     unsigned long long seed, double lo, double hi, hipStream_t s);
// one array of a batched seeded fill (cg_fill_hash_many_f64)
struct HashFill {
  double *d;
  long long sj, sk;
  int nx, ny;
  unsigned long long seed;
  double lo, hi;
};
// f and first (n + 1 prefix counts of points, first[n] = total) in device memory
int launch_fill_hash_many(const HashFill *f, const long long *first, int n, long long total, hipStream_t s);

} // namespace cg
