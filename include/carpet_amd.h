/*
 * carpet_amd.h -- C ABI of libcarpet_amd.so: CarpetLib's inter-grid transfer
 * operators (copy, prolongation, restriction, time interpolation) as
 * hand-written gfx950 (MI355X) HIP kernels.
 *
 * Conventions shared by every entry point
 * ---------------------------------------
 *  - A box ("ibbox3") is 9 ints laid out exactly like CarpetLib's
 *    bbox<int,3>: lower[3], upper[3], stride[3]; bounds are inclusive and
 *    live in the global index space measured in finest-grid units
 *    (CarpetLib/src/bbox.hh:34-70; the reference itself reinterprets an
 *    ibbox3 as int[3][3] at CarpetLib/src/data.cc:321-323).
 *  - An array is a device pointer to a padded 3-D array with x fastest:
 *    element (i,j,k) sits at i + padext[0]*(j + padext[1]*k)
 *    (CarpetLib/src/operator_prototypes_3d.hh:17-28).  ext[] must equal the
 *    array box's sizes() and padext[] >= ext[].
 *  - srcbbox/dstbbox are the array boxes; srcregbbox/dstregbbox the regions.
 *    The operators write dst only inside dstregbbox.
 *  - Every call validates the reference's preconditions on the host before
 *    anything touches the GPU and returns 0 or a negative CG_ERR_* code; the
 *    reference aborts through CCTK_ERROR/CCTK_WARN(0) in the same places.
 *    cg_last_error_message() gives the detail of the last failure of the
 *    calling thread.
 *  - Work is enqueued on the given HIP stream (a hipStream_t passed as
 *    void*; NULL is the default stream) and is not synchronised.
 *  - Floating-point policy (the canonical order): every stencil sum
 *    "res += c*x" of the reference is evaluated left to right starting from
 *    0 as one fused multiply-add per term, res = fma(c, x, res); a
 *    coincident vertex-centred direction passes the value through without
 *    arithmetic; time interpolation is s1*a, then fma(s2,b,.), then
 *    fma(s3,c,.).
 *
 * Entry points that have a reference counterpart cite it; the rest say so.
 */
#ifndef CARPET_AMD_H
#define CARPET_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *cg_stream_t; /* hipStream_t */

/* ---- error codes ------------------------------------------------------ */
enum {
  CG_OK = 0,
  CG_ERR_ARG = -1,          /* null pointer or malformed argument */
  CG_ERR_STRIDE = -2,       /* "Internal error: strides disagree" */
  CG_ERR_REFFACT = -3,      /* strides are not related by a factor 2 */
  CG_ERR_ALIGN = -4,        /* bbox not aligned with vertices / offsets */
  CG_ERR_EMPTY = -5,        /* "region extent is empty" */
  CG_ERR_CONTAIN = -6,      /* region (or stencil footprint) not contained */
  CG_ERR_ORDER = -7,        /* no stencil for this order */
  CG_ERR_TIME_SAME = -8,    /* "arrays have same time" */
  CG_ERR_TIME_EXTRAP = -9,  /* "extrapolation in time" */
  CG_ERR_SHAPE = -10,       /* ext/padext disagree with the array box */
  CG_ERR_OPERATOR = -11,    /* operator/centering combination unsupported */
  CG_ERR_HIP = -12,         /* HIP runtime error */
  CG_ERR_NOMEM = -13,       /* allocation failed */
  CG_ERR_STATE = -14,       /* batch object used in the wrong state */
  CG_ERR_DEVICE = -15,      /* device kernel reported a failed check */
  CG_ERR_SCHEDULE = -16,    /* a dh region check failed (schedule build) */
  CG_ERR_SCRATCH = -17      /* needs more than the allocation-free scratch */
};

/* transport operators: same numbering as CarpetLib's enum operator_type
   (CarpetLib/src/operators.hh:8-25) */
enum {
  CG_OP_ERROR = 0,
  CG_OP_NONE = 1,
  CG_OP_SYNC = 2,
  CG_OP_RESTRICT = 3,
  CG_OP_COPY = 4,
  CG_OP_LAGRANGE = 5,
  CG_OP_ENO = 6,
  CG_OP_WENO = 7,
  CG_OP_TVD = 8,
  CG_OP_LAGRANGE_MONOTONE = 9,
  CG_OP_STAGGER011 = 10,
  CG_OP_STAGGER101 = 11,
  CG_OP_STAGGER110 = 12,
  CG_OP_STAGGER111 = 13
};

/* directions of the staggered operators: bit d set = the field is
   staggered (sits half a coarse spacing off the vertices) in direction d;
   CG_STAGGERxyz matches CarpetLib's op_STAGGERxyz */
enum { CG_STAGGER011 = 6, CG_STAGGER101 = 5, CG_STAGGER110 = 3, CG_STAGGER111 = 7 };

/* centering: same numbering as CarpetLib's enum centering
   (CarpetLib/src/defs.hh:134) */
enum { CG_ERROR_CENTERED = 0, CG_VERTEX_CENTERED = 1, CG_CELL_CENTERED = 2 };

/* precision of Lagrange prolongation in the batch and mixed paths:
   CG_PREC_F64 is the drop-in.  The mixed modes store prolongation sources
   and destinations in FP32 and keep the time interpolation in FP64:
   CG_PREC_MIXED combines the time levels in FP64 as each source window is
   loaded and then runs the stencil once in FP64 (the product of a dyadic
   weight and an FP32 value is exact in FP64 and the sums carry 29 guard
   bits; the result is the FP64 operator on the FP32 values rounded once
   to FP32, to within one FP32 ulp, the difference being the order of the
   time and space sums); CG_PREC_MIXED_DOT2 keeps the levels apart and uses
   compensated FP32 sums (Dot2: FMA TwoProduct + TwoSum per tap) */
enum { CG_PREC_F64 = 0, CG_PREC_MIXED = 1, CG_PREC_MIXED_DOT2 = 2 };

/* Slab description for copies: dst[ipos] = src[ipos + offset*stride].
   Layout of CarpetLib's slab<int,3> (CarpetLib/src/gdata.hh:25-31). */
typedef struct cg_islab {
  int offset[3];
  int is_centered[3];
} cg_islab;

/* A view of one data<T> storage block (CarpetLib/src/data.hh:20-60):
   device storage, padded extents, extents and the array box. */
typedef struct cg_gdata {
  void *storage;
  int padext[3];
  int ext[3];
  int bbox[9];
} cg_gdata;

/* ---- library ---------------------------------------------------------- */
/* no reference counterpart: version and diagnostics */
int cg_version(void);
const char *cg_error_string(int code);
const char *cg_last_error_message(void);
/* no reference counterpart: selects the HIP device for this thread */
int cg_set_device(int device);
int cg_device_count(int *count);

/* ---- device-resident storage ----------------------------------------- */
/* Padded shape of one array: restates pad_shape(shape, offset)
   (CarpetLib/src/cacheinfo.cc:40-209) with VECTORISE_ALIGNED_ARRAYS and
   VECTORISE_ALIGN_FOR_CACHE on and one cache level: x padded to a multiple
   of vec_size, extents shorter than a line rounded to a power of 2, longer
   ones to whole lines; align_interior moves the owned interior (offset) to
   an aligned position.  linesize and vec_size are in elements (linesize a
   power of 2).  Writes padded_shape[3] and padding_offset[3]. */
int cg_pad_shape(const int shape[3], const int offset[3], int vec_size,
                 int linesize, int align_interior, int padded_shape[3],
                 int padding_offset[3]);
/* Device allocation of a padded grid-function block (the device analogue of
   data<T>::allocate, CarpetLib/src/data.cc:355-394 with mem<T>,
   CarpetLib/src/mem.cc:70-140): x-rows padded to 128 bytes and, when
   owned_bbox is given, the owned interior aligned to a row boundary; fills
   out->padext/ext/bbox and out->storage (element (0,0,0)).  Free with
   cg_gdata_free. */
int cg_gdata_alloc(cg_gdata *out, const int bbox[9], const int owned_bbox[9],
                   size_t elem_size);
int cg_gdata_free(cg_gdata *g);
/* no reference counterpart: plain device memory helpers for hosts that have
   no other HIP binding */
int cg_dev_malloc(void **ptr, size_t bytes);
int cg_dev_free(void *ptr);
int cg_memcpy_htod(void *dst, const void *src, size_t bytes);
int cg_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int cg_memcpy_dtod(void *dst, const void *src, size_t bytes, cg_stream_t s);
int cg_memset_d8(void *dst, int value, size_t bytes, cg_stream_t s);
int cg_fill_f64(double *dst, double value, size_t n, cg_stream_t s);
/* no reference counterpart: dst[i] = src[i] for n doubles as a flat stream
   (both arrays 8-byte aligned, else CG_ERR_ARG), the copy ceiling a
   benchmark measures the transfer kernels against */
int cg_stream_copy_f64(double *dst, const double *src, size_t n, cg_stream_t s);
/* no reference counterpart: seeded synthetic data for tests and benchmarks;
   element (i,j,k) of the extent (padding untouched) becomes
   lo + (hi-lo)*u with u the top 53 bits of splitmix64(seed + i +
   ext[0]*(j + ext[1]*k)) scaled to [0,1) */
int cg_fill_hash_f64(double *dst, const int padext[3], const int ext[3],
                     unsigned long long seed, double lo, double hi, cg_stream_t s);
/* no reference counterpart: many cg_fill_hash_f64 fills in one launch
   (setting up thousands of arrays without a launch per array); returns
   once the fills have run on stream s */
typedef struct cg_hash_fill {
  double *dst;       /* element (0,0,0) */
  int padext[3];
  int ext[3];
  unsigned long long seed;
  double lo, hi;
} cg_hash_fill;
int cg_fill_hash_many_f64(int n, const cg_hash_fill *fills, cg_stream_t s);
/* no reference counterpart: smooth synthetic data; every point of g's
   extent at global index X becomes amp * sum_m sin(2 pi k_m.(X*scale) +
   phase_m), nmodes <= 8, kvec[3*m+q] */
int cg_fill_fourier_f64(const cg_gdata *g, double scale, double amp, int nmodes,
                        const int *kvec, const double *phase, cg_stream_t s);
int cg_stream_create(cg_stream_t *s);
int cg_stream_destroy(cg_stream_t s);
int cg_stream_synchronize(cg_stream_t s);
int cg_device_synchronize(void);
/* events on a stream, for timing kernels on the stream they run on */
int cg_event_create(void **ev);
int cg_event_destroy(void *ev);
int cg_event_record(void *ev, cg_stream_t s);
int cg_event_elapsed_ms(float *ms, void *start, void *stop);

/* ---- per-call operators (the plain drop-in path) ---------------------- */
/* copy_3d<double>: CarpetLib/src/copy_3d.cc:25-122. slab may be NULL. */
int cg_copy_3d_f64(const double *src, const int srcpadext[3],
                   const int srcext[3], double *dst, const int dstpadext[3],
                   const int dstext[3], const int srcbbox[9],
                   const int dstbbox[9], const int srcregbbox[9],
                   const int dstregbbox[9], const cg_islab *slab,
                   cg_stream_t stream);

/* prolongate_3d_rf2<double,ORDER>, ORDER in {1,3,5,7,9,11}:
   CarpetLib/src/prolongate_3d_rf2.cc:356-675 */
int cg_prolongate_3d_rf2_f64(int order, const double *src,
                             const int srcpadext[3], const int srcext[3],
                             double *dst, const int dstpadext[3],
                             const int dstext[3], const int srcbbox[9],
                             const int dstbbox[9], const int srcregbbox[9],
                             const int dstregbbox[9], cg_stream_t stream);

/* prolongate_3d_cc_rf2<double,ORDER>, ORDER in 0..5:
   CarpetLib/src/prolongate_3d_cc_rf2.cc:238-579 */
int cg_prolongate_3d_cc_rf2_f64(int order, const double *src,
                                const int srcpadext[3], const int srcext[3],
                                double *dst, const int dstpadext[3],
                                const int dstext[3], const int srcbbox[9],
                                const int dstbbox[9], const int srcregbbox[9],
                                const int dstregbbox[9], cg_stream_t stream);

/* prolongate_3d_eno<CCTK_REAL8> (vertex-centred ENO, refinement factor 2):
   CarpetLib/src/data.cc:201-215 -> prolongate_3d_real8_eno.F90:69-288.
   The dispatch (cg_transfer_from) uses it for vertex-centred op ENO with
   order_space 3 or 5 and op WENO with order_space 5, as data.cc:654-762
   does.  The stencil footprint (coarse points b-1..b+2 around every
   midpoint) must lie in the source array. */
int cg_prolongate_3d_eno_f64(const double *src, const int srcpadext[3], const int srcext[3],
                             double *dst, const int dstpadext[3], const int dstext[3],
                             const int srcbbox[9], const int dstbbox[9], const int srcregbbox[9],
                             const int dstregbbox[9], cg_stream_t stream);

/* prolongate_3d_o5_monotone_rf2<double> (vertex-centred):
   CarpetLib/src/prolongate_3d_o5_monotone_rf2.cc:45-791, the operator of
   op Lagrange_monotone with order_space 5 (data.cc:807-832).  Order-5
   Lagrange summed flat in the reference's term order, replaced by the
   multilinear average where it leaves the range of the nearest coarse
   values.  The footprint (coarse b-2..b+3 on every axis where a point lies
   between coarse points) must lie in the source array. */
int cg_prolongate_3d_o5_monotone_rf2_f64(const double *src, const int srcpadext[3], const int srcext[3],
                                         double *dst, const int dstpadext[3], const int dstext[3],
                                         const int srcbbox[9], const int dstbbox[9],
                                         const int srcregbbox[9], const int dstregbbox[9],
                                         cg_stream_t stream);

/* prolongate_3d_stagger011/101/110/111<double,ORDER>, ORDER in {2,3}:
   CarpetLib/src/prolongate_3d_stagger011.cc:24-561 (the other three
   alike), the operators of op STAGGER011..111 (data.cc:834-920, which maps
   order_space 2 to ORDER 2 and 3..5 to ORDER 3).  stagger: CG_STAGGER*.
   Per direction the tap set depends on staggering and fine parity
   (coeffs_prolongate_3d_stagger.hh); the sum of coeff_x*coeff_y*coeff_z*src
   runs x taps outermost, each operation rounded as written.  Needs
   cg_set_staggered_operators(1) (CarpetLib::support_staggered_operators),
   else CG_ERR_OPERATOR; the tap footprint must lie in the source array. */
int cg_prolongate_3d_stagger_f64(int stagger, int order, const double *src, const int srcpadext[3],
                                 const int srcext[3], double *dst, const int dstpadext[3],
                                 const int dstext[3], const int srcbbox[9], const int dstbbox[9],
                                 const int srcregbbox[9], const int dstregbbox[9], cg_stream_t stream);

/* restrict_3d_stagger011/101/110/111<double>:
   CarpetLib/src/restrict_3d_stagger011.cc:26-153 (the other three alike),
   the restriction of op STAGGER011..111 (data.cc:989-1027): per staggered
   direction the cubic through fine points 2i-1..2i+2 at 2i+1/2
   ({-1,9,9,-1}/16), per unstaggered direction {1,6,1}/8 over 2i-1..2i+1
   (coeffs_restrict_3d_stagger.hh, ORDER_STAG 3).  Needs
   cg_set_staggered_operators(1); the footprint must lie in the source
   array (the reference checks one fine point less on the upper side of
   staggered directions and would read past it). */
int cg_restrict_3d_stagger_f64(int stagger, const double *src, const int srcpadext[3],
                               const int srcext[3], double *dst, const int dstpadext[3],
                               const int dstext[3], const int srcbbox[9], const int dstbbox[9],
                               const int srcregbbox[9], const int dstregbbox[9], cg_stream_t stream);

/* CarpetLib's support_staggered_operators parameter (param.ccl:65,
   default no): the staggered operators return CG_ERR_OPERATOR unless it is
   set.  Process-wide. */
int cg_set_staggered_operators(int support_staggered_operators);

/* prolongate_3d_cc_eno_rf2<double,ORDER>, ORDER in {2,3}:
   CarpetLib/src/prolongate_3d_cc_eno_rf2.cc:943-1363 */
int cg_prolongate_3d_cc_eno_rf2_f64(int order, const double *src,
                                    const int srcpadext[3],
                                    const int srcext[3], double *dst,
                                    const int dstpadext[3],
                                    const int dstext[3], const int srcbbox[9],
                                    const int dstbbox[9],
                                    const int srcregbbox[9],
                                    const int dstregbbox[9],
                                    cg_stream_t stream);

/* prolongate_3d_cc_tvd<CCTK_REAL8> -> prolongate_3d_cc_real8_tvd:
   CarpetLib/src/data.cc:309-324, CarpetLib/src/prolongate_3d_cc_real8_tvd.F90:55-244 */
int cg_prolongate_3d_cc_tvd_f64(const double *src, const int srcpadext[3],
                                const int srcext[3], double *dst,
                                const int dstpadext[3], const int dstext[3],
                                const int srcbbox[9], const int dstbbox[9],
                                const int srcregbbox[9],
                                const int dstregbbox[9], cg_stream_t stream);

/* restrict_3d_rf2<double>: CarpetLib/src/restrict_3d_rf2.cc:25-121 */
int cg_restrict_3d_rf2_f64(const double *src, const int srcpadext[3],
                           const int srcext[3], double *dst,
                           const int dstpadext[3], const int dstext[3],
                           const int srcbbox[9], const int dstbbox[9],
                           const int srcregbbox[9], const int dstregbbox[9],
                           cg_stream_t stream);

/* restrict_3d_cc_rf2<double>: CarpetLib/src/restrict_3d_cc_rf2.cc:23-152 */
int cg_restrict_3d_cc_rf2_f64(const double *src, const int srcpadext[3],
                              const int srcext[3], double *dst,
                              const int dstpadext[3], const int dstext[3],
                              const int srcbbox[9], const int dstbbox[9],
                              const int srcregbbox[9],
                              const int dstregbbox[9], cg_stream_t stream);

/* interpolate_3d_2tl<double>: CarpetLib/src/interpolate_3d_2tl.cc:24-117 */
int cg_interpolate_3d_2tl_f64(const double *src1, double t1,
                              const double *src2, double t2,
                              const int srcpadext[3], const int srcext[3],
                              double *dst, double t, const int dstpadext[3],
                              const int dstext[3], const int srcbbox[9],
                              const int dstbbox[9], const int srcregbbox[9],
                              const int dstregbbox[9], cg_stream_t stream);

/* interpolate_3d_4tl<double> / interpolate_3d_5tl<double>:
   CarpetLib/src/interpolate_3d_4tl.cc:24-131, interpolate_3d_5tl.cc:24-140
   (cubic / quartic Lagrange in time, data.cc:1227-1256); the sum in the
   canonical order s1*src1, then fma(s_l, src_l, .) */
int cg_interpolate_3d_4tl_f64(const double *src1, double t1, const double *src2, double t2,
                              const double *src3, double t3, const double *src4, double t4,
                              const int srcpadext[3], const int srcext[3], double *dst, double t,
                              const int dstpadext[3], const int dstext[3], const int srcbbox[9],
                              const int dstbbox[9], const int srcregbbox[9], const int dstregbbox[9],
                              cg_stream_t stream);
int cg_interpolate_3d_5tl_f64(const double *src1, double t1, const double *src2, double t2,
                              const double *src3, double t3, const double *src4, double t4,
                              const double *src5, double t5, const int srcpadext[3], const int srcext[3],
                              double *dst, double t, const int dstpadext[3], const int dstext[3],
                              const int srcbbox[9], const int dstbbox[9], const int srcregbbox[9],
                              const int dstregbbox[9], cg_stream_t stream);

/* interpolate_3d_3tl<double>: CarpetLib/src/interpolate_3d_3tl.cc:24-126 */
int cg_interpolate_3d_3tl_f64(const double *src1, double t1,
                              const double *src2, double t2,
                              const double *src3, double t3,
                              const int srcpadext[3], const int srcext[3],
                              double *dst, double t, const int dstpadext[3],
                              const int dstext[3], const int srcbbox[9],
                              const int dstbbox[9], const int srcregbbox[9],
                              const int dstregbbox[9], cg_stream_t stream);

/* restrict_3d_cc_o3_rf2<double> / restrict_3d_cc_o5_rf2<double>:
   CarpetLib/src/restrict_3d_cc_o3_rf2.cc, restrict_3d_cc_o5_rf2.cc -- the
   cubic (quintic) through 4 (6) fine cells per axis evaluated at the coarse
   cell centre (not conservative); the footprint, fine cells 2i-1..2i+2
   (2i-2..2i+3) per axis, must lie in the source array or CG_ERR_CONTAIN
   is returned */
int cg_restrict_3d_cc_o3_rf2_f64(const double *src, const int srcpadext[3],
                                 const int srcext[3], double *dst,
                                 const int dstpadext[3], const int dstext[3],
                                 const int srcbbox[9], const int dstbbox[9],
                                 const int srcregbbox[9], const int dstregbbox[9],
                                 cg_stream_t stream);
int cg_restrict_3d_cc_o5_rf2_f64(const double *src, const int srcpadext[3],
                                 const int srcext[3], double *dst,
                                 const int dstpadext[3], const int dstext[3],
                                 const int srcbbox[9], const int dstbbox[9],
                                 const int srcregbbox[9], const int dstregbbox[9],
                                 cg_stream_t stream);
/* restrict_3d_vc_rf2<double, centi, centj, centk>:
   CarpetLib/src/restrict_3d_vc_rf2.cc:26-260 -- the restriction of a
   cell-centred group's slab with one vertex-centred direction (is_centered
   (0,1,1), (1,0,1) or (1,1,0), data.cc:1106-1127; face-centred data such
   as refluxing fluxes).  Per coarse point nested half-sums, x innermost: a
   cell-centred direction (cent = 1) averages fine points srcoff+2i and
   srcoff+2i+1, a vertex-centred one takes srcoff+2i, with srcoff = (reg.lo
   - src.lo - src.stride/2) / src.stride; every point read must lie in the
   source array.  cg_transfer_from dispatches it for a cell-centred
   restriction whose slab has such an is_centered. */
int cg_restrict_3d_vc_rf2_f64(int centi, int centj, int centk, const double *src,
                              const int srcpadext[3], const int srcext[3], double *dst,
                              const int dstpadext[3], const int dstext[3], const int srcbbox[9],
                              const int dstbbox[9], const int srcregbbox[9], const int dstregbbox[9],
                              cg_stream_t stream);
/* CarpetLib's use_dgfe parameter (param.ccl:30, default no).  When set, the
   reference replaces cell-centred Lagrange prolongation and full-cell
   restriction by prolongate_3d_dgfe_rf2 / restrict_3d_dgfe_rf2<T,5>
   (data.cc:617-627, 1055-1062), which evaluate the discontinuous-Galerkin
   element operators of the optional HRSCCore thorn (hrscc.hh,
   GLLElement<5>); that code is not part of the reference tree and this
   library does not provide it, so with the parameter set those transfers
   return CG_ERR_OPERATOR instead of silently running Lagrange.
   Process-wide. */
int cg_set_dgfe(int use_dgfe);
/* CarpetLib's parameters use_higher_order_restriction and
   restriction_order_space (CarpetLib/param.ccl) as the transfer dispatch
   sees them: cell-centred restrictions of groups whose operator is not
   ENO/WENO use the o3 or o5 stencil when enabled (CarpetLib/src/data.cc,
   transfer_restrict).  Process-wide; defaults as param.ccl: off, order 3;
   an order other than 1, 3, 5 returns CG_ERR_ORDER */
int cg_set_restriction(int use_higher_order_restriction, int restriction_order_space);

/* interpolate_eno_3d_3tl<double>: CarpetLib/src/interpolate_eno_3d_3tl.cc:
   88-164 -- quadratic in time; a point where that makes a new extremum
   falls back to linear interpolation between the two levels whose interval
   holds t (the time interpolation of ENO/TVD groups, data.cc:1282-1292) */
int cg_interpolate_eno_3d_3tl_f64(const double *src1, double t1,
                                  const double *src2, double t2,
                                  const double *src3, double t3,
                                  const int srcpadext[3], const int srcext[3],
                                  double *dst, double t, const int dstpadext[3],
                                  const int dstext[3], const int srcbbox[9],
                                  const int dstbbox[9], const int srcregbbox[9],
                                  const int dstregbbox[9], cg_stream_t stream);

/* ---- mixed precision (no reference counterpart: the FP32 form beside the
   FP64 drop-in).  Sources and destination hold float; every 1-D sum runs in
   FP32 with compensated (TwoProduct/TwoSum, Dot2) accumulation; the fused
   time interpolation runs in FP64 and the result is rounded to FP32 once.
   Same arguments and validation as the FP64 entry points. */
int cg_prolongate_3d_rf2_mixed(int order, const float *src, const int srcpadext[3],
                               const int srcext[3], float *dst, const int dstpadext[3],
                               const int dstext[3], const int srcbbox[9],
                               const int dstbbox[9], const int srcregbbox[9],
                               const int dstregbbox[9], cg_stream_t stream);
int cg_prolongate_3d_cc_rf2_mixed(int order, const float *src, const int srcpadext[3],
                                  const int srcext[3], float *dst, const int dstpadext[3],
                                  const int dstext[3], const int srcbbox[9],
                                  const int dstbbox[9], const int srcregbbox[9],
                                  const int dstregbbox[9], cg_stream_t stream);
/* cg_transfer_from for float storage: Lagrange prolongation only (with
   time interpolation in FP64); copies, restrictions, TVD and ENO stay FP64
   and are rejected here (CG_ERR_OPERATOR) */
int cg_transfer_from_mixed(const cg_gdata *dst, const cg_gdata *const *srcs,
                           const double *times, int ntimes, const int dstregbbox[9],
                           const int srcregbbox[9], double time, int transport_operator,
                           int centering, int order_space, int order_time,
                           cg_stream_t stream);

/* ---- transfer dispatch ------------------------------------------------ */
/* data<T>::transfer_from_innerloop -> transfer_time -> transfer_p_r ->
   {copy_from_innerloop, transfer_prolongate, transfer_restrict} and
   time_interpolate, with gdata::find_source_timelevel's time-level choice
   (CarpetLib/src/data.cc:464-1314, CarpetLib/src/gdata.cc:288-349).
   srcs[0..ntimes-1] are the source time levels (all with the same box),
   times[] their times.  Where the reference prolongates every time level
   into a temporary and then interpolates in time, this call runs one fused
   kernel that keeps the prolongated values in registers; the arithmetic is
   the same.  With 4 or 5 time levels (order_time 3 or 4; Lagrange, copy
   and staggered groups only, as in data.cc:1192-1258) the levels go
   through region-sized temporaries from a stream-ordered allocation, as in
   the reference; cg_batch_add refuses such prolongations.  Cell-centred
   restrictions follow cg_set_restriction. */
int cg_transfer_from(const cg_gdata *dst, const cg_gdata *const *srcs,
                     const double *times, int ntimes,
                     const int dstregbbox[9], const int srcregbbox[9],
                     const cg_islab *slab, double time, int transport_operator,
                     int centering, int order_space, int order_time,
                     cg_stream_t stream);

/* No reference counterpart: cg_transfer_from for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving
   loops, hosts that forbid hipMalloc).  Same arguments, validation and
   results as cg_transfer_from.  The transfers that cg_transfer_from runs
   through stream-ordered temporaries (prolongation with 4 or 5 time levels,
   restriction and slabbed copies with time interpolation) stage them here
   in one library-owned device array per GPU of CG_NOALLOC_SCRATCH_BYTES,
   reserved when the library is loaded on that GPU and held until the
   process ends; a transfer whose temporaries (time levels x region points
   x 8 bytes) exceed it returns CG_ERR_SCRATCH before any GPU work.  Every
   other transfer is one kernel on the caller's arrays.  Nothing is
   allocated or freed (no hipMalloc, hipMallocAsync or hipFree).  All calls
   on one GPU share its array: calls whose work overlaps on one GPU (on
   different streams, or from several threads) race on it and must be
   ordered by the caller.  No part of the array needs clearing: every
   temporary point is written before it is read. */
#define CG_NOALLOC_SCRATCH_BYTES (64ull << 20)
int cg_transfer_from_noalloc(const cg_gdata *dst, const cg_gdata *const *srcs,
                             const double *times, int ntimes,
                             const int dstregbbox[9], const int srcregbbox[9],
                             const cg_islab *slab, double time, int transport_operator,
                             int centering, int order_space, int order_time,
                             cg_stream_t stream);

/* ---- batched level step ----------------------------------------------- */
/* Replaces the per-region walk of ggf::transfer_from_all
   (CarpetLib/src/ggf.cc:500-591) for a whole level step: the caller adds
   every (region, variable) transfer once (phases order the dependent
   steps, e.g. restrict=0, coarse sync=1, prolongate=2, fine sync=3), the
   batch builds a device-resident descriptor list, and each (phase, kernel
   class) runs as one launch over it -- one workgroup per work item, or a
   persistent grid taking items from per-XCD counters; the whole step can
   be captured as one hipGraph and replayed while the hierarchy is
   unchanged, with new target times and rotated time levels between
   replays (cg_batch_set_time, cg_batch_set_storage). */
typedef struct cg_batch cg_batch;
int cg_batch_create(cg_batch **out);
int cg_batch_destroy(cg_batch *b);
/* same arguments as cg_transfer_from plus the phase; returns the transfer
   index (>= 0) or an error code.  prec selects CG_PREC_F64, CG_PREC_MIXED
   or CG_PREC_MIXED_DOT2 (mixed only for Lagrange prolongation; dst/src storage
   then holds float). */
int cg_batch_add(cg_batch *b, int phase, const cg_gdata *dst,
                 const cg_gdata *const *srcs, const double *times, int ntimes,
                 const int dstregbbox[9], const int srcregbbox[9],
                 const cg_islab *slab, double time, int transport_operator,
                 int centering, int order_space, int order_time, int prec);
/* builds tiles and uploads the descriptor list; tile_cfg < 0 lets the
   tuner choose */
int cg_batch_finalize(cg_batch *b, int tile_cfg);
/* runs every phase, one launch per (phase, kernel class) */
int cg_batch_run(cg_batch *b, cg_stream_t stream);
/* runs the same work as one launch per (region, variable) transfer: the
   "plain" per-box baseline */
int cg_batch_run_plain(cg_batch *b, cg_stream_t stream);
/* captures cg_batch_run into a hipGraph, then replays it */
int cg_batch_capture_graph(cg_batch *b, cg_stream_t stream);
int cg_batch_launch_graph(cg_batch *b, cg_stream_t stream);
/* Per-step parameters of a batch: what changes between the reference's
   level steps without a regrid -- the target time of each substep, which
   ggf::transfer_from_all passes down to data<T>::transfer_time and
   gdata::find_source_timelevel (CarpetLib/src/ggf.cc:500-591,
   data.cc:485-534, gdata.cc:288-349), and the rotation of the time levels'
   storage by ggf::cycle_all (CarpetLib/src/ggf.cc:248-265).  Each call
   resolves the affected transfers again exactly as cg_batch_add did (same
   validation, time-level choice and weights) and keeps the launches, work
   items and a captured graph: the changed descriptors are uploaded in
   stream order by the next cg_batch_run or cg_batch_launch_graph on that
   stream (a pinned staging copy; nothing inside the graph changes).  A
   transfer must stay in its launch class: a tensor prolongation built with
   2 or 3 time levels may also resolve to 1 (a target time that matches a
   level exactly runs in the same class on that level alone, bit-identical
   to a one-level transfer); one built at an exact match (1 level) cannot
   later interpolate; operators and stencils cannot change.  Otherwise
   CG_ERR_STATE, and the batch is unchanged.
   cg_batch_set_time: every transfer of `phase` added with exactly ntimes
   source levels gets these level times and target time; returns the number
   of transfers updated.
   cg_batch_set_storage: every destination or source level of any transfer
   whose storage is from[i].storage now uses to[i] (the same layout:
   padext, ext, bbox), all pairs applied at once -- ggf::cycle_all is the
   rotation from = {tl0, tl1, tl2}, to = {tl2, tl0, tl1} per grid function;
   returns the number of transfers updated. */
int cg_batch_set_time(cg_batch *b, int phase, const double *times, int ntimes, double time);
int cg_batch_set_storage(cg_batch *b, int n, const cg_gdata *from, const cg_gdata *to);
/* Autotuner: times each launch of the finalized batch under candidate
   work-item sizes (tensor tiles: z chunk of 8/16/32/64/128/all coarse bases;
   pointwise items: 256, 512, 1024, 2048, 4096 or 16384 points) in the
   class's default mode
   (persistent grid for tensor tiles, one workgroup per item for pointwise
   runs), then the best size in the other modes (one workgroup per item,
   persistent grids of resident capacity and of 2 workgroups per CU) and,
   for tensor tiles, the other tile shape (wide: 32x8 coarse bases, 64x8
   for CG_PREC_MIXED; narrow: 16x16),
   with HIP events, and keeps the fastest per launch class.  FP64 regions
   thin in x and in y or z (shell edges and corners) form classes of their
   own on 8x8-base tiles, and regions at most 4 coarse bases across in x
   and long in y (shell faces normal to x) on 4x16-base tiles; these two
   shapes are fixed.
   Choices are keyed on (kernel class, order, time levels, precision, log2
   of the median region size, log2 of the region count) and kept in the
   text file cache_path (may be NULL; lines "<key>|<zchunk> <points>
   <workgroups per CU: -1 one per item, 0 resident capacity> <tile shape:
   0 wide, 1 narrow, 2 small, -1 by region size> <256-point items per
   workgroup in grid mode: 1, 2, 4, 8, 16>"): classes the file names take its
   parameters, the others are searched (for pointwise classes of 256-point
   items in grid mode also the items per workgroup).  Each trial builds its
   launch alone into its own descriptor list.  Then every phase of several
   launches is timed with them side by side and one after the other (see
   cg_batch_set_phase_mode) and keeps the faster, cached as lines
   "phase <key> + <key> ...|<1 side by side, 0 one after the other>".  The
   search executes each launch (one phase's class of transfers) reps times
   on its own, outside phase order, so call it before the step whose
   results are kept, or on scratch data.  Drops a captured graph.  With the
   environment variable CG_TUNE_LOG set, every trial's time goes to
   stderr. */
int cg_batch_autotune(cg_batch *b, cg_stream_t stream, int reps,
                      const char *cache_path);
/* no reference counterpart: how the launches of one phase run.  They write
   disjoint regions, so they may run side by side (helper streams forked
   from the caller's stream, joined before the next phase; graph branches
   under capture) or one after the other.  mode 1 side by side, 0 one after
   the other, -1 automatic: the tuner's choice, else side by side exactly
   when every launch of the phase runs one workgroup per item (a persistent
   launch fills the GPU by itself).  Any time; applies from the next
   cg_batch_run or capture and drops a captured graph. */
int cg_batch_set_phase_mode(cg_batch *b, int phase, int mode);
/* human-readable summary: a totals line (transfers, launches, work items,
   bytes of the device descriptor blob), then one line per launch (phase,
   class key, work items, grid, tile choice, grid mode; in phases of several
   launches "side by side" or "in turn"); returns the full length */
int cg_batch_describe(const cg_batch *b, char *buf, size_t len);
/* no reference counterpart: the launches of a finalized batch one at a
   time, for per-kernel timing -- their number, launch idx run on its own
   (its work counters cleared first), and its phase, destination points and
   algorithmic bytes */
int cg_batch_launch_count(const cg_batch *b);
int cg_batch_run_launch(cg_batch *b, int idx, cg_stream_t stream);
int cg_batch_launch_stats(const cg_batch *b, int idx, int *phase, long long *dst_points,
                          long long *alg_bytes);
/* statistics: number of transfers, destination points and algorithmic
   bytes (destination + unique source footprint per time level) */
int cg_batch_stats(const cg_batch *b, long long *ntransfers,
                   long long *dst_points, long long *alg_bytes);

/* ---- communication schedule ------------------------------------------
   dh::regrid's region description and communication lists for one
   multigrid level (CarpetLib/src/dh.cc:211-1080): per level the components'
   exterior / owned / interior / communicated boxes, the buffer / overlap /
   active split, and the lists
     CG_SCHED_REF_PROL      refinement prolongation, fast_ref_prol_sendrecv
                            (dh.cc:766-830): level rl-1 -> active + overlaps
     CG_SCHED_SYNC          synchronisation, fast_sync_sendrecv
                            (dh.cc:834-890): interiors -> ghost zones
     CG_SCHED_REF_BND_PROL  boundary prolongation, fast_ref_bnd_prol_sendrecv
                            (dh.cc:894-975): level rl-1 -> unsynchronised
                            ghosts + buffer zones
     CG_SCHED_REF_REST      refinement restriction, fast_ref_rest_sendrecv
                            (dh.cc:981-1080), stored on the coarse level like
                            the reference's: level rl+1 -> level rl
   Each entry is (source component, send box, destination component,
   receive box): one cg_transfer_from (or cg_batch_add) with srcregbbox =
   send and dstregbbox = recv.  Region sets may be split into boxes
   differently from the reference's bboxset normal form; each list's union
   of receive boxes per (source, destination) is the same, and so is every
   transferred value.  Refluxing and multigrid lists are not built.
   The reference's consistency assertions return CG_ERR_SCHEDULE with the
   reference's message in cg_last_error_message. */
typedef struct cg_schedule cg_schedule;
typedef struct cg_level_spec {
  int baseextent[9];        /* h.baseextent(ml, rl): the level's domain exterior */
  int ncomp;                /* components of the level */
  const int *interior;      /* ncomp x 9: h.extent(ml, rl, c), the regridder's boxes */
  int reffact[3];           /* h.reffacts[rl] (level 0 = 1) */
  int ghost_width[6];       /* lower x y z, upper x y z */
  int buffer_width[6];
  int overlap_width[6];
  int prolongation_order;   /* prolongation_orders_space[rl]; stencil = order / 2 */
} cg_level_spec;
enum { CG_SCHED_REF_PROL = 0, CG_SCHED_SYNC = 1, CG_SCHED_REF_BND_PROL = 2, CG_SCHED_REF_REST = 3,
       CG_SCHED_OLD2NEW_SYNC = 4, CG_SCHED_OLD2NEW_PROL = 5, CG_SCHED_NLISTS = 6 };
/* boxes of a component (cg_schedule_component) */
enum { CG_SCHED_EXTERIOR = 0, CG_SCHED_INTERIOR = 1, CG_SCHED_OWNED = 2, CG_SCHED_COMMUNICATED = 3 };
/* region sets of a component (cg_schedule_set) */
enum { CG_SCHED_GHOSTS = 10, CG_SCHED_BUFFERS = 11, CG_SCHED_OVERLAPS = 12, CG_SCHED_ACTIVE = 13,
       CG_SCHED_SYNCED = 14, CG_SCHED_BNDREF = 15, CG_SCHED_BOUNDARIES = 16 };
/* centering: CG_VERTEX_CENTERED or CG_CELL_CENTERED (h.refcent);
   boundary_width: h.boundary_width (lower x y z, upper x y z);
   use_higher_order_restriction / restriction_order_space: CarpetLib's
   parameters (restriction into interiors only, and for cell centring a
   source shrunk by order/2, dh.cc:1019-1050).  The process-wide
   support_staggered_operators (cg_set_staggered_operators) applies too, as
   in the reference: restriction into interiors only, sources shrunk by 2. */
int cg_schedule_build(int centering, const int boundary_width[6], int nlevels,
                      const cg_level_spec *levels, int use_higher_order_restriction,
                      int restriction_order_space, cg_schedule **out);
void cg_schedule_free(cg_schedule *s);
/* entries of list on level rl, or a negative CG_ERR_* */
int cg_schedule_count(const cg_schedule *s, int rl, int list);
int cg_schedule_region(const cg_schedule *s, int rl, int list, int i, int *src_comp, int send[9],
                       int *dst_comp, int recv[9]);
int cg_schedule_component(const cg_schedule *s, int rl, int c, int which, int bbox[9]);
/* boxes of a region set: writes min(n, cap) boxes (9 ints each) when bboxes
   is not NULL and returns n */
int cg_schedule_set(const cg_schedule *s, int rl, int c, int which, int cap, int *bboxes);
/* regridding schedule of s after a regrid from old (dh.cc:1712-1850, the
   lists ggf::recompose_fill walks, ggf.cc:175-215): for every component of
   s, its active and overlap points from the owned regions of old's
   components on the same level (CG_SCHED_OLD2NEW_SYNC, send = recv, the
   source component indexes old), the rest by prolongation from s's next
   coarser level (CG_SCHED_OLD2NEW_PROL).  old may be NULL (initial data:
   prolongation only, nothing asserted); with an old hierarchy every point
   must be filled (CG_ERR_SCHEDULE otherwise).  Replaces earlier regrid
   lists of s. */
int cg_schedule_regrid(cg_schedule *s, const cg_schedule *old);

#ifdef __cplusplus
}
#endif

#endif /* CARPET_AMD_H */
