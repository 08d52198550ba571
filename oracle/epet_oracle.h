/*
 * epet_oracle.h -- CPU oracle for the epetreco MLEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load this library, and only as the
 * checker / the timed CPU restatement.  The product (libepet.so) never links,
 * loads or calls it.
 *
 * Every function restates one piece of the reference algorithm in plain C,
 * single-precision where the reference is single-precision, with the same
 * order of floating-point operations, so that results are bit-identical with
 * the reference's host arithmetic (gcc, SSE, no FMA contraction).  Each
 * function cites the reference file:line it follows (paths relative to the
 * reference's src/ directory).
 *
 * Parity pins (what checks each function against something other than itself)
 * are listed beside each declaration; see tests/test_oracle_*.py.
 */
#ifndef EPET_ORACLE_H
#define EPET_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field order of DefaultMeasurementSetup<float> (MeasurementSetup.hpp:129-139). */
typedef struct {
  float pos0x, pos1x;
  int   na;
  float da;
  int   n0z, n0y, n1z, n1y;
  float segx, segy, segz;
} eo_setup;

/* Field order of DefaultVoxelGrid<float> (VoxelGrid.hpp:88-91): O[3], D[3], N[3]. */
typedef struct {
  float o[3];
  float d[3];
  int   n[3];
} eo_grid;

/* ---- index maps -------------------------------------------------------- */
/* MeasurementSetupLinIndex.hpp:32-165.  Pinned: closed-form round trip over
 * every channel id (reference test test_MeasurementSetupLinIndex.cu:37-46). */
int eo_ms_linid(const eo_setup* s, int id0z, int id0y, int id1z, int id1y, int ida);
int eo_ms_ida (const eo_setup* s, int lin);
int eo_ms_id0z(const eo_setup* s, int lin);
int eo_ms_id0y(const eo_setup* s, int lin);
int eo_ms_id1z(const eo_setup* s, int lin);
int eo_ms_id1y(const eo_setup* s, int lin);
/* VoxelGridLinIndex.hpp:193-282 (z fastest).  Pinned: round trip
 * (test_VoxelGridLinIndex.cu:37-44). */
int eo_vg_linid(const eo_grid* g, int idx, int idy, int idz);
int eo_vg_idx(const eo_grid* g, int lin);
int eo_vg_idy(const eo_grid* g, int lin);
int eo_vg_idz(const eo_grid* g, int lin);

/* ---- geometry ---------------------------------------------------------- */
/* sinf/cosf of the rotation angle, argument evaluated as
 * ((float(M_PI)*ida)*da)/180.f (MeasurementSetupTrafo2CartCoord.hpp:67-68). */
void eo_angle_sincos(const eo_setup* s, int ida, float* sin_, float* cos_);
/* Box coordinates b in [0,1]^3 of a segment of detector `det` (0 or 1) to
 * world coordinates (MeasurementSetupTrafo2CartCoord.hpp:52-105 for det 0,
 * :221-274 for det 1).  Pinned: bit-exact against the host-compiled reference
 * functors (oracle/_ref) and against the closed form R_y(theta)(pos+(b-1/2)e). */
void eo_trafo(const eo_setup* s, int det, const float box[3],
              int idz, int idy, int ida, float out[3]);
/* distancePointLine.h:19-55.  Pinned: _ref bit-exact + double brute force. */
float eo_absolute(const float a[3]);
void  eo_cross(float r[3], const float a[3], const float b[3]);
float eo_distance(const float a[3], const float b[3], const float p[3]);

/* ---- workqueue (candidate structure) ------------------------------------ */
/* Candidate predicate of getWorkqueueEntry (getWorkqueue.hpp:50-78).
 * Pinned: _ref bit-exact on full config-1 lists and sampled config-2 rows. */
int eo_is_candidate(const eo_setup* s, const eo_grid* g, int cnl, int vxl);
/* getWorkqueue with cursor (listId, vxlId) and optional cap (cap<0: none)
 * (getWorkqueue.hpp:129-223).  Writes up to cap pairs when the output arrays
 * are non-NULL; returns the number found and advances the cursor exactly as
 * the reference.  Pinned: _ref bit-exact; prefix invariant
 * (test_getWorkqueue.cu:186-198). */
long long eo_workqueue(const eo_setup* s, const eo_grid* g,
                       const int* list, int list_size,
                       int* listId, int* vxlId, long long cap,
                       int* out_cnl, int* out_vxl);

/* ---- Monte-Carlo rays --------------------------------------------------- */
/* Philox4x32-10 (Salmon et al., SC'11).  Pinned: Random123 known-answer
 * vectors and the open-source rocRAND header implementation. */
void eo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Uniform in (0,1): odd multiples of 2^-24, exactly representable. */
float eo_u01(uint32_t x);
/* Ray r of channel cnl: counters {r, cnl, 0, 0} and {r, cnl, 1, 0}, key =
 * seed; 6 uniforms -> trafo of detector 0 (ray[0..2]) and 1 (ray[3..5]), the
 * role RandRayGen plays in RayGenerators.hpp:29-35. */
void eo_ray(const eo_setup* s, int cnl, int r, uint64_t seed, float ray[6]);
/* Deterministic pixel-centre ray (ReguRayGen, RayGenerators.hpp:42-66). */
void eo_ray_center(const eo_setup* s, int cnl, float ray[6]);

/* ---- system-matrix values ------------------------------------------------ */
/* Slab intersection length of the infinite line through ray with a voxel box
 * (getSystemMatrixDeviceOnly.cu:88-133).  Pinned: closed forms (axis ray = d,
 * diagonal = d*sqrt3), double-precision brute-force sampling. */
float eo_intersection_length(const float vxl[6], const float ray[6]);
/* Voxel box corners (getSystemMatrixDeviceOnly.cu:163-175). */
void  eo_voxel_box(const eo_grid* g, int vxl, float box[6]);
/* sum_r intersectionLength(vxl, ray_r) / nrays with the per-channel rays of
 * eo_ray (getSystemMatrixFromWorkqueue.cu:143-256 model; calcSme
 * getSystemMatrixDeviceOnly.cu:189-201 summation order). */
float eo_calc_sme(const eo_setup* s, const eo_grid* g, int cnl, int vxl,
                  int nrays, uint64_t seed);
/* Same, with the rays given (nrays x 6 floats). */
float eo_calc_sme_rays(const eo_grid* g, int vxl, const float* rays, int nrays);
/* Values for a workqueue (getSystemMatrixFromWorkqueue.cu:97-261). */
void  eo_sm_from_workqueue(const eo_setup* s, const eo_grid* g,
                           const int* cnl, const int* vxl, float* val,
                           long long n, int nrays, uint64_t seed);
/* Row-major system matrix of the listed channels: every candidate voxel of
 * every row (keep_zeros=1, the reference's getSystemMatrix output after
 * cooSort/COO->CSR/CSR->ECSR, wrappers.hpp:172-247) or only candidates with a
 * nonzero value (keep_zeros=0).  rowptr has m+1 entries.  Returns nnz, or -1
 * if cap is exceeded.  Per-row counts only when col/val are NULL. */
long long eo_sm_rows(const eo_setup* s, const eo_grid* g, const int* list, int m,
                     int nrays, uint64_t seed, int keep_zeros,
                     int64_t* rowptr, int* col, float* val, long long cap);

/* ---- Siddon chords -------------------------------------------------------- */
/* getChords (ChordsCalc_lowlevel.hpp:173-357), with a capacity check the
 * reference lacks.  Returns the number of chords written (>=0), or -1 when
 * nChords is too small.  Pinned: sum of chords = length of the ray inside the
 * grid (double), hand-checked 2x2x2 case. */
int eo_get_chords(float* chords, int* voxelIds, int nChords, const float ray[6],
                  const float gridO[3], const float gridD[3], const int gridN[3]);

/* ---- sparse formats ------------------------------------------------------ */
/* cooSort (cooSort.hpp:53-71): lexicographic (row, col) order, stable.
 * Pinned: reference known answer test_cooSort.cu:17-19, std::sort. */
void eo_coo_sort(float* val, int* row, int* col, long long n);
/* convertCoo2Csr (cusparseWrapper.hpp:38-45).  Pinned: test_cusparseWrapper.cu:26-28. */
void eo_coo2csr(int* csrRowPtr, const int* cooRowId, long long nnz, int m);
/* convertCsr2Ecsr (convertCsr2Ecsr.hpp:34-47).  Pinned: test_convertCsr2Ecsr.cu:7-8. */
void eo_csr2ecsr(int* ecsr, const int* yRowId, int nnzr, const int* csr, int m);

/* ---- algebra -------------------------------------------------------------- */
/* CSRmv<float> semantics y = alpha*op(A)*x + beta*y (csrmv.hpp:27-35),
 * trans 0 = N, 1 = T, accumulation in double.  Pinned: dense product,
 * adjoint identity <Ax,y> = <x,A^T y>. */
void eo_csrmv(int trans, int m, int n, float alpha, const float* val,
              const int64_t* rowptr, const int* col, const float* x,
              float beta, float* y);
/* mlemOperations.hpp:22-139.  Pinned: test_mlemOperations.cu:39-95 answers. */
void  eo_divides(float* res, const float* a, const float* b, long long n);
void  eo_divides_multiplies(float* res, const float* a, const float* b,
                            const float* c, long long n);
void  eo_scales(float* x, float alpha, long long n);
float eo_sum(const float* x, long long n);
double eo_sum_f64(const float* x, long long n);

/* ---- MLEM ------------------------------------------------------------------ */
/* The iteration loop of tests/reco.cu:209-300 (single rank, one chunk) on an
 * ECSR matrix (m rows = measured channels, n = voxels).  x is the initial
 * image on input (the caller normalises it as reco.cu:129-137 does) and the
 * result on output; s is the sensitivity, normalised here to sum 1
 * (reco.cu:182-183).  eo_mlem_f32 follows the reference's FP32 operations
 * (unguarded y/yhat); eo_mlem_f64 is the FP64 oracle (ratio 0 where yhat=0).
 * Pinned: A = I gives y/sum(y) after one iteration; sum x = 1; nonnegativity;
 * Poisson log-likelihood nondecreasing; noise-free convergence. */
void eo_mlem_f32(int m, int n, const float* val, const int64_t* rowptr,
                 const int* col, const float* y, const float* s, float* x,
                 int niter);
void eo_mlem_f64(int m, int n, const float* val, const int64_t* rowptr,
                 const int* col, const float* y, const float* s, double* x,
                 int niter);
/* F frames of eo_mlem_f64 sharing one matrix, frame-minor (X[j*F+f],
 * Y[i*F+f]); frame f runs niter_f[f] iterations and is then left unchanged.
 * The back projection walks an internal transpose whose rows keep ascending
 * channel order, so each frame's summation order, and hence its result, is
 * bit-identical to eo_mlem_f64 on that frame alone, for any thread count. */
void eo_mlem_f64_frames(int m, int n, int F, const float* val, const int64_t* rowptr,
                        const int* col, const float* Y, const float* s, double* X,
                        const int* niter_f, int nthreads);
/* eo_mlem_f64_frames for matrices of billions of entries: the back projection
 * scatters into one private FP64 image per thread (static row ranges) and
 * sums them in thread order, so no transpose is built.  Deterministic for a
 * given thread count; equal to eo_mlem_f64_frames up to FP64 summation order
 * (pinned against it at 1e-12). */
void eo_mlem_f64_frames_par(int m, int n, int F, const float* val, const int64_t* rowptr,
                            const int* col, const float* Y, const float* s, double* X,
                            const int* niter_f, int nthreads);
/* Sensitivity image s_j = sum_i a_ij accumulated in FP64: the back projection
 * of an all-ones measurement (tests/backprojection.cu:115-161) that
 * tests/reco.cu:150-183 reads as its sensitivity.  Per-thread private images
 * over static row ranges, summed in thread order. */
void eo_colsum_f64(int m, int n, const float* val, const int64_t* rowptr, const int* col,
                   double* s, int nthreads);
/* One FP32 MLEM iteration with OpenMP (the CPU baseline).  Same result as
 * one step of eo_mlem_f32 up to summation order; yhat/e/c scratch provided. */
void eo_mlem_iter_omp(int m, int n, const float* val, const int64_t* rowptr,
                      const int* col, const float* y, const float* s, float* x,
                      float* yhat, float* c, int nthreads);
/* Poisson log-likelihood sum_i y_i log(yhat_i) - yhat_i (double). */
double eo_loglik(int m, const float* val, const int64_t* rowptr, const int* col,
                 const float* y, const double* x);

/* ---- chunk helpers (wrappers.hpp:405-445) --------------------------------- */
uint32_t eo_nInChunk(uint32_t idChunk, uint32_t nAll, uint32_t nFullChunk);
uint32_t eo_chunkPtr(uint32_t chunkId, uint32_t nFullChunk);
uint32_t eo_nChunks(uint64_t maxNnz, uint64_t maxNChunk);

/* Number of OpenMP threads the library would use. */
int eo_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
