// cg_kernels.hip -- per-call launches of the pointwise kernels, the tensor
// launch dispatch (instances in cg_tensor_{vlag,clag,eno}.hip) and the
// synthetic-data fills (bodies in cg_tiles.hpp).

#include <hip/hip_runtime.h>

#include "cg_tiles.hpp"

namespace cg {

long long tensor_tiles(TensorArgs &a, const TileCfg &cfg) {
  a.ntiles[0] = (a.nq[0] + cfg.cx - 1) / cfg.cx;
  a.ntiles[1] = (a.nq[1] + cfg.cy - 1) / cfg.cy;
  a.zchunk = cfg.zchunk > 0 ? cfg.zchunk : a.nq[2];
  a.ntiles[2] = (a.nq[2] + a.zchunk - 1) / a.zchunk;
  return (long long)a.ntiles[0] * a.ntiles[1] * a.ntiles[2];
}

int launch_tensor(int kind, int order, int ntl, int prec, const TensorArgs &a,
                  const TileCfg &cfg, hipStream_t s) {
  if (prec < 0 || prec > 2 || (prec != 0 && fp64_only(kind)))
    return -1;
  switch (kind) {
  case LK_VLAG: return launch_tensor_kind<LK_VLAG>(order, ntl, prec, a, cfg, s);
  case LK_CLAG: return launch_tensor_kind<LK_CLAG>(order, ntl, prec, a, cfg, s);
  case LK_ENO: return launch_tensor_kind<LK_ENO>(order, ntl, prec, a, cfg, s);
  case LK_VENO: return launch_tensor_kind<LK_VENO>(order, ntl, prec, a, cfg, s);
  }
  return -1;
}

template <int OP>
__global__ __launch_bounds__(256) void k_point(PointArgs a) {
  const unsigned nx = (unsigned)a.n[0];
  const unsigned nxy = nx * (unsigned)a.n[1];
  const unsigned e = blockIdx.x * 256u + threadIdx.x;
  if (e >= nxy)
    return;
  const int j = (int)(e / nx);
  const int i = (int)(e - (unsigned)j * nx);
  for (int k = blockIdx.y; k < a.n[2]; k += gridDim.y)
    point_elem<OP>(a, i, j, k);
}

// TVD: one thread per coarse cell of the region's parents, x-y plane
// flattened, z strided over blockIdx.y
template <int NTL>
__global__ __launch_bounds__(256) void k_tvd(PointArgs a) {
  int qx0, qy0, qz0, nqx, nqy, nqz;
  tvd_range(a, 0, &qx0, &nqx);
  tvd_range(a, 1, &qy0, &nqy);
  tvd_range(a, 2, &qz0, &nqz);
  const unsigned e = blockIdx.x * 256u + threadIdx.x;
  if (e >= (unsigned)nqx * (unsigned)nqy)
    return;
  const int cy = (int)(e / (unsigned)nqx);
  const int cx = (int)(e - (unsigned)cy * (unsigned)nqx);
  for (int cz = blockIdx.y; cz < nqz; cz += gridDim.y)
    tvd_cell<NTL>(a, qx0 + cx, qy0 + cy, qz0 + cz);
}

int launch_point(int op, const PointArgs &a, hipStream_t s) {
  if (op == PO_TVD) {
    // coarse parents per axis: ((off + n - 1) >> 1) - (off >> 1) + 1
    long long nq[3];
    for (int d = 0; d < 3; ++d)
      nq[d] = ((a.off[d] + a.n[d] - 1) >> 1) - (a.off[d] >> 1) + 1;
    const long long nxy = nq[0] * nq[1];
    if (nxy <= 0 || nq[2] <= 0 || nxy >= (1LL << 31))
      return -1;
    dim3 grid((unsigned)((nxy + 255) / 256), (unsigned)(nq[2] < 65535 ? nq[2] : 65535));
    switch (a.ntl) {
    case 1: hipLaunchKernelGGL(k_tvd<1>, grid, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_tvd<2>, grid, dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_tvd<3>, grid, dim3(256), 0, s, a); break;
    default: return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  const long long nxy = (long long)a.n[0] * a.n[1];
  if (nxy <= 0 || a.n[2] <= 0 || nxy >= (1LL << 32) / 2)
    return -1;
  const unsigned gx = (unsigned)((nxy + 255) / 256);
  const unsigned gy = (unsigned)(a.n[2] < 65535 ? a.n[2] : 65535);
  dim3 grid(gx, gy), block(256);
  switch (op) {
  case PO_COPY: hipLaunchKernelGGL(k_point<PO_COPY>, grid, block, 0, s, a); break;
  case PO_RESTRICT_VC: hipLaunchKernelGGL(k_point<PO_RESTRICT_VC>, grid, block, 0, s, a); break;
  case PO_RESTRICT_CC: hipLaunchKernelGGL(k_point<PO_RESTRICT_CC>, grid, block, 0, s, a); break;
  case PO_INTERP: hipLaunchKernelGGL(k_point<PO_INTERP>, grid, block, 0, s, a); break;
  case PO_RESTRICT_CC_O3: hipLaunchKernelGGL(k_point<PO_RESTRICT_CC_O3>, grid, block, 0, s, a); break;
  case PO_RESTRICT_CC_O5: hipLaunchKernelGGL(k_point<PO_RESTRICT_CC_O5>, grid, block, 0, s, a); break;
  case PO_MONO5: hipLaunchKernelGGL(k_point<PO_MONO5>, grid, block, 0, s, a); break;
  case PO_STAGGER: hipLaunchKernelGGL(k_point<PO_STAGGER>, grid, block, 0, s, a); break;
  case PO_RESTRICT_STAGGER: hipLaunchKernelGGL(k_point<PO_RESTRICT_STAGGER>, grid, block, 0, s, a); break;
  case PO_RESTRICT_VC2: hipLaunchKernelGGL(k_point<PO_RESTRICT_VC2>, grid, block, 0, s, a); break;
  default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

__global__ void k_fill(double *d, double v, size_t n) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t st = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += st)
    d[i] = v;
}

// splitmix64 counter stream: element (i,j,k) of the extent gets
// lo + (hi-lo) * u(seed + i + ext0*(j + ext1*k)), u in [0,1) from the top 53 bits
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  unsigned long long z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ void k_fill_hash(double *d, long long sj, long long sk, int nx, int ny, int nz,
                            unsigned long long seed, double lo, double hi) {
  const long long n = (long long)nx * ny * nz;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const long long k = e / ((long long)nx * ny);
    const long long r = e - k * nx * ny;
    const long long j = r / nx, i = r - j * nx;
    const unsigned long long z = splitmix64(seed + (unsigned long long)e);
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    d[i + sj * j + sk * k] = lo + (hi - lo) * u;
  }
}

// Many seeded fills in one launch: array a owns the global point range
// [first[a], first[a+1]) (x fastest within its extent); workgroup b fills
// the FILL_CHUNK points from b * FILL_CHUNK, starting at the last array
// whose range begins at or before them.  Same values as k_fill_hash.
constexpr long long FILL_CHUNK = 8192;

__global__ void k_fill_hash_many(const HashFill *__restrict__ f, const long long *__restrict__ first, int n,
                                 long long total) {
  const long long c0 = (long long)blockIdx.x * FILL_CHUNK;
  const long long c1 = min(total, c0 + FILL_CHUNK);
  __shared__ int s_a;
  if (threadIdx.x == 0) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (first[mid] <= c0)
        lo = mid;
      else
        hi = mid - 1;
    }
    s_a = lo;
  }
  __syncthreads();
  int a = s_a;
  for (long long e = c0 + threadIdx.x; e < c1; e += blockDim.x) {
    while (e >= first[a + 1]) // e < total = first[n], so a stays below n
      ++a;
    const HashFill &h = f[a];
    const long long l = e - first[a];
    const long long nxy = (long long)h.nx * h.ny;
    const long long k = l / nxy, r = l - k * nxy;
    const long long j = r / h.nx, i = r - j * h.nx;
    const unsigned long long z = splitmix64(h.seed + (unsigned long long)l);
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    h.d[i + h.sj * j + h.sk * k] = h.lo + (h.hi - h.lo) * u;
  }
}

int launch_fill_hash_many(const HashFill *f, const long long *first, int n, long long total, hipStream_t s) {
  if (total <= 0 || n <= 0)
    return 0;
  const long long blocks = (total + FILL_CHUNK - 1) / FILL_CHUNK;
  if (blocks > 0x7fffffffLL)
    return -1;
  hipLaunchKernelGGL(k_fill_hash_many, dim3((unsigned)blocks), dim3(256), 0, s, f, first, n, total);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_fill_hash(double *d, long long sj, long long sk, const int *ext,
                     unsigned long long seed, double lo, double hi, hipStream_t s) {
  const long long n = (long long)ext[0] * ext[1] * ext[2];
  if (n <= 0)
    return 0;
  long long blocks = (n + 255) / 256;
  if (blocks > 65536)
    blocks = 65536;
  hipLaunchKernelGGL(k_fill_hash, dim3((unsigned)blocks), dim3(256), 0, s, d, sj, sk, ext[0],
                     ext[1], ext[2], seed, lo, hi);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// sum of Fourier modes: amp * sum_m sin(2 pi (k_m . x) + phase_m) at
// x = (lower + idx*stride) * scale for every point of the extent
struct FourierArgs {
  double *d;
  long long sj, sk;
  int n[3];
  int lower[3], stride[3];
  double scale, amp;
  int nmodes;
  int k[8][3];
  double phase[8];
};

__global__ void k_fill_fourier(FourierArgs a) {
  const long long n = (long long)a.n[0] * a.n[1] * a.n[2];
  const double twopi = 6.283185307179586476925286766559;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const long long k = e / ((long long)a.n[0] * a.n[1]);
    const long long r = e - k * a.n[0] * a.n[1];
    const long long j = r / a.n[0], i = r - j * a.n[0];
    const double x = (a.lower[0] + i * a.stride[0]) * a.scale;
    const double y = (a.lower[1] + j * a.stride[1]) * a.scale;
    const double z = (a.lower[2] + k * a.stride[2]) * a.scale;
    double v = 0.0;
    for (int m = 0; m < a.nmodes; ++m)
      v += sin(twopi * (a.k[m][0] * x + a.k[m][1] * y + a.k[m][2] * z) + a.phase[m]);
    a.d[i + a.sj * j + a.sk * k] = a.amp * v;
  }
}

int launch_fill_fourier(double *d, long long sj, long long sk, const int *ext,
                        const int *lower, const int *stride, double scale, double amp,
                        int nmodes, const int *kvec, const double *phase, hipStream_t s) {
  if (nmodes < 0 || nmodes > 8)
    return -1;
  FourierArgs a{};
  a.d = d;
  a.sj = sj;
  a.sk = sk;
  for (int q = 0; q < 3; ++q) {
    a.n[q] = ext[q];
    a.lower[q] = lower[q];
    a.stride[q] = stride[q];
  }
  a.scale = scale;
  a.amp = amp;
  a.nmodes = nmodes;
  for (int m = 0; m < nmodes; ++m) {
    for (int q = 0; q < 3; ++q)
      a.k[m][q] = kvec[3 * m + q];
    a.phase[m] = phase[m];
  }
  const long long n = (long long)ext[0] * ext[1] * ext[2];
  if (n <= 0)
    return 0;
  long long blocks = (n + 255) / 256;
  if (blocks > 65536)
    blocks = 65536;
  hipLaunchKernelGGL(k_fill_fourier, dim3((unsigned)blocks), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

// flat copy of n doubles: two per thread, 256 apart, both loads before the
// stores.  Measured on 2 x 2.1 GB: 6.45 TB/s read + written, against 6.35
// with one double per thread, 6.0 with four and 5.65-6.15 with 16-byte
// accesses (one to four per thread, plain or nontemporal).
__global__ __launch_bounds__(256) void k_stream_copy(const double *__restrict__ src,
                                                     double *__restrict__ dst, size_t n) {
  const size_t b = (size_t)blockIdx.x * 512 + threadIdx.x;
  double v0 = 0, v1 = 0;
  if (b < n)
    v0 = src[b];
  if (b + 256 < n)
    v1 = src[b + 256];
  if (b < n)
    dst[b] = v0;
  if (b + 256 < n)
    dst[b + 256] = v1;
}

int launch_stream_copy(double *dst, const double *src, size_t n, hipStream_t s) {
  if (n == 0)
    return 0;
  const size_t blocks = (n + 511) / 512;
  if (blocks > 0x7fffffff)
    return -1;
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int launch_fill(double *dst, double v, size_t n, hipStream_t s) {
  if (n == 0)
    return 0;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536)
    blocks = 65536;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, dst, v, n);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

} // namespace cg
