/*
 * epet.h -- C-ABI of the MI355X-native epetreco MLEM hot path (libepet.so).
 *
 * Plain C types only: device buffers are raw device pointers (hipMalloc,
 * torch, or epet_malloc), sizes are integers, streams are opaque `void*`
 * hipStream_t values.  Every function returns EPET_OK (0) or an error code and
 * leaves a message in epet_last_error() (thread-local); the C++ surface in
 * include/epetreco/ turns error codes into exceptions, keeping the reference's
 * "throw on device error" contract (CUDA_HandleError.hpp:17-31).
 *
 * All work is enqueued on the context's stream.  Calls that return host values
 * (counts, sums, cursors) synchronise that stream, as the reference's calls do
 * (wrappers.hpp:211, mlemOperations.hpp:138).  A context is not thread-safe.
 *
 * Reference citations are relative to the reference's src/ directory.
 */
#ifndef EPET_H
#define EPET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes -------------------------------------------------------- */
#define EPET_OK            0
#define EPET_ERR_ARG       1   /* invalid argument (null, negative size, ...) */
#define EPET_ERR_HIP       2   /* HIP runtime error */
#define EPET_ERR_CAPACITY  3   /* output buffer too small */
#define EPET_ERR_STATE     4   /* geometry not set, wrong matrix type, ... */
#define EPET_ERR_ALLOC     5   /* device allocation failed */
#define EPET_ERR_UNSUPPORTED 6 /* size or mode outside the implemented range */

/* value storage of a library-owned matrix */
#define EPET_VAL_F32   0
#define EPET_VAL_BF16  1   /* bf16 value x per-row FP32 scale */
#define EPET_VAL_F16   2   /* fp16 value x per-row FP32 scale */

/* operation flag of CSRmv (cusparseOperation_t role, csrmv.hpp:17-22) */
#define EPET_OP_N  0
#define EPET_OP_T  1

/* limits */
#define EPET_MAX_ANGLES 256   /* per-angle sin/cos table size */
#define EPET_MAX_RAYS   256   /* rays per channel */

/* Measurement setup; member layout of DefaultMeasurementSetup<float>
 * (MeasurementSetup.hpp:129-139).  Units: metres, degrees. */
typedef struct {
  float pos0x, pos1x;
  int   na;
  float da;
  int   n0z, n0y, n1z, n1y;
  float segx, segy, segz;
} epet_setup;

/* Voxel grid; member layout of DefaultVoxelGrid<float> (VoxelGrid.hpp:88-91). */
typedef struct {
  float o[3];
  float d[3];
  int   n[3];
} epet_grid;

typedef struct epet_ctx epet_ctx;
typedef struct epet_matrix epet_matrix;

/* ---- errors, contexts ------------------------------------------------------ */
const char* epet_last_error(void);
int epet_version(void);

/* Context on `device`, launching on `stream` (NULL = the null stream).
 * Replaces the reference's implicit default device/stream and its cuSPARSE
 * handle (tests/reco.cu:102-105). */
int epet_ctx_create(epet_ctx** out, int device, void* stream);
int epet_ctx_destroy(epet_ctx* ctx);
int epet_ctx_set_stream(epet_ctx* ctx, void* stream);
int epet_ctx_synchronize(epet_ctx* ctx);

/* Geometry, rays per channel and Philox seed of the context.  Replaces the
 * cudaMemcpyToSymbol uploads of setup_const / grid_const / nrays_const
 * (device_constant_memory.hpp:14-26, tests/reco.cu:64-72).  The per-angle
 * sin/cos table is computed here on the host, with the reference's argument
 * order (MeasurementSetupTrafo2CartCoord.hpp:67-68), so device geometry is
 * bit-identical with host geometry.  Kernels receive it in the kernel-argument
 * (constant) segment. */
int epet_set_geometry(epet_ctx* ctx, const epet_setup* setup, const epet_grid* grid);
int epet_set_nrays(epet_ctx* ctx, int nrays);
int epet_set_seed(epet_ctx* ctx, uint64_t seed);

/* ---- device memory (cuda_wrappers.hpp:15-87 role) -------------------------- */
int epet_malloc(epet_ctx* ctx, void** ptr, size_t bytes);
int epet_free(epet_ctx* ctx, void* ptr);
int epet_memcpy_h2d(epet_ctx* ctx, void* dst, const void* src, size_t bytes);
int epet_memcpy_d2h(epet_ctx* ctx, void* dst, const void* src, size_t bytes);
int epet_memcpy_d2d(epet_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- Monte-Carlo rays (replaces RandRayGen, RayGenerators.hpp:11-40) -------- */
/* rays[(i*nrays + r)*6 + 0..5] = start (detector 0) and end (detector 1) of ray
 * r of channel cnl[i]: Philox4x32-10 with counters {r, cnl, 0|1, 0} and key =
 * seed, six uniforms (odd multiples of 2^-24) through the segment transform.
 * Host replay: identical bits. */
int epet_rays(epet_ctx* ctx, const int* cnl_dev, int n, float* rays_dev);

/* Siddon chords of a batch of rays (getChords, ChordsCalc_lowlevel.hpp:173-357)
 * on the context's grid: rays_dev holds nrays x 6 floats (start, end).  Ray r
 * writes its chords to chords_dev[r*nChords + k] and the voxel (ix, iy, iz)
 * of each chord to voxelIds_dev[3*(r*nChords + k)]; counts_dev[r] = chords
 * written, or -1 when nChords is too small (a check the reference lacks).
 * The reference's quirks are kept: a ray with an endpoint inside the grid
 * yields no chords (:208-218), non-crossed axes take part in the crossing
 * test (:312), the first alpha step starts at aDimmin + aDimup (:242). */
int epet_getChords(epet_ctx* ctx, const float* rays_dev, int nrays, int nChords,
                   float* chords_dev, int* voxelIds_dev, int* counts_dev);

/* ---- workqueue (getWorkqueue.hpp:129-223) ----------------------------------- */
/* Scans (list entry, voxel) pairs from the cursor (*listId, *vxlId) in
 * lexicographic order and emits candidate pairs (getWorkqueue.hpp:50-78),
 * at most `cap` of them (cap < 0: all; then out_capacity must hold them).
 * Advances the cursor exactly as the reference.  *n_found = pairs written.
 * With both output pointers NULL and out_capacity 0 it only counts: *n_found =
 * the number of pairs the call would emit, cursor unchanged. */
int epet_getWorkqueue(epet_ctx* ctx, const int* list_dev, int list_size,
                      int* listId, int* vxlId, long long cap,
                      int* out_cnl_dev, int* out_vxl_dev, long long out_capacity,
                      long long* n_found);

/* ---- system matrix ---------------------------------------------------------- */
/* getSystemMatrixFromWorkqueue (getSystemMatrixFromWorkqueue.cu:97-102):
 * val[i] = sum_r intersectionLength(voxel vxl[i], ray r of cnl[i]) / nrays. */
int epet_getSystemMatrixFromWorkqueue(epet_ctx* ctx, const int* cnl_dev,
                                      const int* vxl_dev, float* val_dev,
                                      long long wqLength, int nrays);

/* getSystemMatrix (getSystemMatrixDeviceOnly.cu:324-331): COO triplets of
 * the `mlSize` listed channels.  The reference writes every candidate, zeros
 * included; keep_zeros = 1 keeps that structure, 0 keeps nonzeros only.
 * Entries are appended at *truckDest_dev (a device u64 count, as
 * truckDest_devi) and come out in row-major order.  `capacity` = length of
 * the output arrays. */
int epet_getSystemMatrix(epet_ctx* ctx, float* sme_dev, int* vxlId_dev,
                         int* cnlId_dev, const int* ml_dev, uint32_t mlSize,
                         unsigned long long* truckDest_dev, long long capacity,
                         int keep_zeros);

/* systemMatrixCalculation (wrappers.hpp:172-177): getSystemMatrix ->
 * cooSort -> COO->CSR over `nchannels` rows -> CSR->ECSR for the nY listed
 * rows.  Buffers are the caller's, as in the reference (wrappers.hpp:64-74):
 * ecsrPtr[nY+1], vxlId/val/cnlId[capacity], csrPtr[nchannels+1].  *nnz_dev
 * (device u64) receives the entry count.  keep_zeros as above. */
int epet_systemMatrixCalculation(epet_ctx* ctx, int* ecsrPtr_dev, int* vxlId_dev,
                                 float* val_dev, unsigned long long* nnz_dev,
                                 int* cnlId_dev, int* csrPtr_dev,
                                 const int* yRowId_dev, uint32_t nY,
                                 int nchannels, long long capacity, int keep_zeros);

/* Fused build of the nonzero system matrix of m listed channels into a
 * library-owned ECSR matrix (int64 row pointers, int32 voxel ids in ascending
 * order, FP32 values): the replacement of the reference's per-iteration
 * rebuild (tests/reco.cu:225).  One workgroup per channel: Philox rays, ray
 * traversal marking the voxels each ray may cross, the reference's exact slab
 * value for every marked voxel, the reference's candidate predicate, ordered
 * ballot/prefix-sum compaction.  Bit-identical with the reference algorithm
 * given the same rays.  list_dev may be NULL: rows 0..m-1 are channels 0..m-1. */
int epet_sm_build(epet_ctx* ctx, const int* list_dev, int m, epet_matrix** out);

int epet_matrix_info(const epet_matrix* A, int* m, int* n, long long* nnz,
                     int* value_type);
/* Device arrays of a matrix (rowscale NULL for FP32 storage). */
int epet_matrix_arrays(const epet_matrix* A, int64_t** rowptr, int** col,
                       void** val, float** rowscale);
int epet_matrix_destroy(epet_matrix* A);
/* Library matrix from caller arrays (copied): rowptr int64[m+1], col, FP32 val. */
int epet_matrix_create(epet_ctx* ctx, int m, int n, long long nnz,
                       const int64_t* rowptr_dev, const int* col_dev,
                       const float* val_dev, epet_matrix** out);
/* Rows rows_dev[0..nrows) of A as a new matrix (the ECSR restriction of a
 * full-channel matrix to the measured channels, convertCsr2Ecsr.hpp:34-47). */
int epet_matrix_select_rows(epet_ctx* ctx, const epet_matrix* A, const int* rows_dev,
                            int nrows, epet_matrix** out);
/* Copy of A with values stored as value_type (EPET_VAL_BF16 / EPET_VAL_F16:
 * per-row scale = row max, stored value = a / scale rounded to nearest even). */
int epet_matrix_convert(epet_ctx* ctx, const epet_matrix* A, int value_type,
                        epet_matrix** out);
/* Attach a transposed copy of A (rows = voxels, entries in ascending row
 * order, built deterministically without atomics); the multi-frame back
 * projection of epet_mlem then gathers from it.  Needs strictly ascending
 * columns within rows. */
int epet_matrix_transpose(epet_ctx* ctx, epet_matrix* A);
/* Column sums in FP64 (the sensitivity image s_j = sum_i a_ij). */
int epet_matrix_colsum_f64(epet_ctx* ctx, const epet_matrix* A, double* s_dev);

/* ---- sparse format conversion ---------------------------------------------- */
/* cooSort (cooSort.hpp:53-59): in-place, stable, lexicographic (row, col). */
int epet_cooSort(epet_ctx* ctx, float* val_dev, int* row_dev, int* col_dev, long long N);
/* convertCoo2Csr (cusparseWrapper.hpp:38-42): zero-based row pointers of m
 * rows from nnz row ids. */
int epet_convertCoo2Csr(epet_ctx* ctx, int* csrRowPtr_dev, const int* cooRowId_dev,
                        long long nnz, int m);
/* convertCsr2Ecsr (convertCsr2Ecsr.hpp:34-39): ecsr[i] = csr[yRowId[i]],
 * ecsr[nnzr] = csr[m]. */
int epet_convertCsr2Ecsr(epet_ctx* ctx, int* ecsr_dev, const int* yRowId_dev,
                         int nnzr, const int* csr_dev, int m);

/* ---- sparse products --------------------------------------------------------- */
/* CSRmv<float> (csrmv.hpp:27-35): y = alpha*op(A)*x + beta*y, int32 row
 * pointers, general zero-based matrix (cusparseWrapper.hpp:21-26).  op N: x has
 * n entries, y m; op T: x has m, y n.  alpha/beta are host pointers. */
int epet_csrmv(epet_ctx* ctx, int op, int m, int n, long long nnz,
               const float* alpha, const float* val_dev, const int* rowptr_dev,
               const int* col_dev, const float* x_dev, const float* beta,
               float* y_dev);
/* Same product on a library matrix (any value storage). */
int epet_matrix_spmv(epet_ctx* ctx, const epet_matrix* A, int op, float alpha,
                     const float* x_dev, float beta, float* y_dev);
/* Multi-vector product: X and Y are frame-minor (X[j*nrhs + f]), nrhs <= 64.
 * op N: Y(m x nrhs) = alpha*A*X + beta*Y; op T: Y(n x nrhs) = alpha*A^T X + beta*Y. */
int epet_matrix_spmm(epet_ctx* ctx, const epet_matrix* A, int op, int nrhs,
                     float alpha, const float* X_dev, float beta, float* Y_dev);

/* ---- MLEM element-wise operations (mlemOperations.hpp:22-139) ----------------- */
int epet_divides(epet_ctx* ctx, float* res, const float* a, const float* b, long long n);
int epet_dividesMultiplies(epet_ctx* ctx, float* res, const float* a, const float* b,
                           const float* c, long long n);
int epet_scales(epet_ctx* ctx, float* x, float alpha, long long n);
/* Sum with FP64 accumulation (the reference's FP32 thrust::reduce role);
 * synchronises and returns the value on the host. */
int epet_sum(epet_ctx* ctx, const float* x, long long n, double* out);

/* ---- fused MLEM (tests/reco.cu:209-300 loop) ----------------------------------- */
/* niter MLEM iterations on device for nframes frames sharing A:
 *   yhat = A x;  e = y / yhat (0 where yhat = 0);  c = A^T e;
 *   x <- x * c / (s + [s==0]);  x <- x / sum(x)
 * y: m x nframes, x: n x nframes (frame-minor), s: n (normalised to sum 1
 * inside).  x holds the initial image on entry (the caller normalises it, as
 * reco.cu:129-137 does).  loglik_out (host, niter*nframes doubles, may be
 * NULL) receives the FP64 Poisson log-likelihood sum_i y log yhat - yhat of the
 * image entering each iteration.  With use_graph = 1 one iteration is captured
 * as a hipGraph and replayed.  frame_active (host, nframes ints, may be NULL)
 * stops a frame once its relative log-likelihood change is below tol (tol <= 0:
 * never); it receives the number of iterations each frame ran. */
int epet_mlem(epet_ctx* ctx, const epet_matrix* A, int nframes, const float* y_dev,
              const float* s_dev, float* x_dev, int niter, int use_graph,
              double tol, double* loglik_out, int* frame_iters);

/* Single-frame MLEM that allocates nothing: the same iteration as epet_mlem
 * with nframes = 1, use_graph = 0, tol = 0, run eagerly on the context's
 * stream, with its working arrays (the ratio e: m floats; the back projection
 * c and the normalised s: n floats each; niter FP64 log-likelihoods) in a
 * fixed device array of EPET_NOALLOC_SCRATCH_BYTES per GPU that the library
 * holds until the process ends.  It makes no hipMalloc/hipFree call and
 * builds no product plan: it runs the plans A already holds (those that an
 * earlier epet_mlem call on A built) and otherwise the plan-free kernels.
 * Sizes whose working arrays exceed the limit fail with EPET_ERR_CAPACITY
 * before any work.  Calls share their GPU's array and each clears what it
 * needs: calls that overlap on one GPU (two host threads, or two streams)
 * race.  loglik_out (host, niter doubles) may be NULL.  Returns after the
 * stream has finished the iterations (it synchronises, so it is not for use
 * under stream capture).  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (tight serving loops, harnesses that
 * forbid allocation). */
#define EPET_NOALLOC_SCRATCH_BYTES (64ull << 20)
int epet_mlem_noalloc(epet_ctx* ctx, const epet_matrix* A, const float* y_dev,
                      const float* s_dev, float* x_dev, int niter, double* loglik_out);

/* Autotuner of the multi-frame products epet_mlem runs for F frames on A and
 * its attached transpose.  Blocked products (the default): every shape of
 * (rows per block rb in {128,256,512,1024}) x (frames per lane P in {4,8}) that
 * fits the workgroup's LDS is timed per direction and the fastest kept;
 * cfg = {rb_fwd, P_fwd, rb_bwd, P_bwd}.  Gather kernels
 * (EPET_BATCHED=gather, or blocked copies that cannot be allocated): (rows per
 * wavefront R in {1,2,4,8}) x (frames per lane tile P in {4,8,16,32});
 * cfg = {R_fwd, P_fwd, R_bwd, P_bwd}.  cfg (4 ints) and ms (2 floats: forward,
 * back projection) may be NULL.  epet_mlem runs it on first use for a frame
 * count.  No reference counterpart: the reference's loop is single-frame
 * (tests/reco.cu:232-250). */
int epet_mlem_tune(epet_ctx* ctx, epet_matrix* A, int F, int* cfg, float* ms);

/* The multi-frame path epet_mlem runs for F frames on A, as tuned: info (10
 * ints).  Blocked products: {2, P_fwd, lanes per row slot fwd, rows per block
 * fwd, splits fwd, P_bwd, lanes bwd, rows per block bwd, splits bwd, 0}.
 * Gather kernels: {1, P_fwd, R_fwd, 0, 0, P_bwd, R_bwd, 0, 0, 0}.  {0, ...}:
 * not tuned yet.  ms (2 floats, may be NULL) = the tuned times per launch. */
int epet_mlem_kernels(epet_ctx* ctx, const epet_matrix* A, int F, int* info, float* ms);

/* ---- chunk helpers (wrappers.hpp:405-445) ---------------------------------------- */
uint32_t epet_nInChunk(uint32_t idChunk, uint32_t nAll, uint32_t nFullChunk);
uint32_t epet_chunkPtr(uint32_t chunkId, uint32_t nFullChunk);
uint32_t epet_nChunks(unsigned long long maxNnz, unsigned long long maxNChunk);

#ifdef __cplusplus
}
#endif
#endif /* EPET_H */
