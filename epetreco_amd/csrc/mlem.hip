// mlem.hip -- mlemOperations drop-ins and the fused MLEM iteration.
//
// One MLEM iteration (tests/reco.cu:209-300) is four launches on the context's
// stream, replayable as one hipGraph (no memset nodes; every clear is a
// kernel):
//   1. k_fwd_ratio : yhat_i = sum_k a_ik x_k, e_i = y_i / yhat_i (0 where
//                    yhat_i = 0), FP64 log-likelihood sum y log yhat - yhat
//   2. back projection c += A^T e (spmv.hip kernels)
//   3. k_update    : x_j <- x_j c_j / (s_j + [s_j == 0]), FP64 sum of x
//   4. k_normalise : x_j <- x_j / sum x, c <- 0, next sum slot <- 0, it++
// Device state (iteration counter, sums, log-likelihoods, frame flags) lives in
// one small buffer so that the same graph serves every iteration.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "epet_block.hpp"
#include "epet_internal.hpp"
#include "epet_rows.hpp"

namespace epet {

int matrix_spmv(epet_ctx* ctx, const epet_matrix* A, int op, float alpha, const float* x,
                float beta, float* y);
int matrix_spmm(epet_ctx* ctx, const epet_matrix* A, int op, int F, float alpha, const float* X,
                float beta, float* Y);
int pick_lanes(long long nnz, int m);
int bp_prepare(epet_ctx* ctx, epet_matrix* A);
bool bp_use_transpose(const epet_ctx* ctx, const epet_matrix* A);
int matrix_bp_transpose(epet_ctx* ctx, const epet_matrix* At, float alpha, const float* x,
                        float beta, float* y);
int matrix_spmv_t_prepared(epet_ctx* ctx, const epet_matrix* A, float alpha, const float* x,
                           float beta, float* y);
int lds_prepare(epet_ctx* ctx, epet_matrix* A);
int lds_prepare_map(epet_ctx* ctx, epet_matrix* A, bool voxel_columns);
bool lds_ready(const epet_ctx* ctx, const epet_matrix* A);
int lds_fwd_ratio(epet_ctx* ctx, const epet_matrix* A, const float* x, const float* y, float* e,
                  int fold, const int* itp, double* loglik);
int mf_tune(epet_ctx* ctx, epet_matrix* A, const MfArgs& a, int reps);
int mf_iteration(epet_ctx* ctx, const epet_matrix* A, const MfArgs& a);
int blk_prepare(epet_ctx* ctx, epet_matrix* A, int F);
bool blk_ready(const epet_matrix* A, int F);
bool blk_built(const epet_matrix* A, int F);
int blk_tune(epet_ctx* ctx, epet_matrix* A, const MfArgs& a, int reps);
int blk_iteration(epet_ctx* ctx, const epet_matrix* A, const MfArgs& a);
int blk_info(const epet_matrix* A, int F, int info[9], float ms[2]);

// whether multi-frame MLEM on A runs the blocked products (copies built and
// tuned for F frames, not turned off by the context)
static bool use_blk(const epet_ctx* ctx, const epet_matrix* A, int F) {
  return !ctx->opt.batched_gather && blk_ready(A, F);
}

// ------------------------------------------------------------------ element-wise
__global__ __launch_bounds__(256) void k_divides(float* __restrict__ res, const float* __restrict__ a,
                                                 const float* __restrict__ b, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) res[i] = a[i] / b[i];   // mlemOperations.hpp:22-32, unguarded
}

__global__ __launch_bounds__(256) void k_divides_multiplies(float* __restrict__ res,
                                                            const float* __restrict__ a,
                                                            const float* __restrict__ b,
                                                            const float* __restrict__ c,
                                                            long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    float dvs = c[i];
    dvs += (float)(dvs == 0.0f);   // mlemOperations.hpp:48-50
    res[i] = a[i] * b[i] / dvs;
  }
}

__global__ __launch_bounds__(256) void k_scales(float* __restrict__ x, float alpha, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = alpha * x[i];
}

__device__ __forceinline__ double block_sum_f64(double v, double* s_w) {
  v = wave_sum(v);
  if (lane_id() == 0) s_w[wave_id()] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x / WAVE); w++) t += s_w[w];
  return t;   // valid in thread 0
}

__global__ __launch_bounds__(256) void k_sum_f64(const float* __restrict__ x, long long n,
                                                 double* __restrict__ out) {
  __shared__ double s_w[256 / WAVE];
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    acc += x[i];
  const double t = block_sum_f64(acc, s_w);
  if (threadIdx.x == 0) unsafeAtomicAdd(out, t);
}

__global__ void k_zero_f64(double* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ------------------------------------------------------------------ fused MLEM
struct MlemState {
  int it;          // iterations completed
  int F;
  double* sum;     // [2][F]
  double* loglik;  // [niter][F]
  int* active;     // [F]
  int niter;
};


// Single frame, short rows: L lanes per row, grid-stride over rows; one FP64
// log-likelihood atomic per workgroup.  fold = 1 multiplies the ratio by the row
// scale, for a back projection through the transposed copy (which stores the
// unscaled values).
template <typename VT, int L>
__global__ __launch_bounds__(256) void k_fwd_ratio(int m, const int64_t* __restrict__ rowptr,
                                                   const int* __restrict__ col,
                                                   const typename VT::T* __restrict__ val,
                                                   const float* __restrict__ rowscale,
                                                   const float* __restrict__ x,
                                                   const float* __restrict__ y,
                                                   float* __restrict__ e, int fold,
                                                   const int* __restrict__ itp,
                                                   double* __restrict__ loglik) {
  __shared__ double s_w[256 / WAVE];
  const int sub = threadIdx.x % L;
  const long long stride = (long long)gridDim.x * (256 / L);
  double ll = 0.0;
  for (long long r = (long long)blockIdx.x * (256 / L) + threadIdx.x / L; r < m; r += stride) {
    const long long b = rowptr[r], end = rowptr[r + 1];
    float a0 = 0.0f, a1 = 0.0f;
    long long k = b + sub;
    for (; k + L < end; k += 2 * L) {
      a0 += VT::get(val, k) * x[ldcol(col, k)];
      a1 += VT::get(val, k + L) * x[ldcol(col, k + L)];
    }
    if (k < end) a0 += VT::get(val, k) * x[ldcol(col, k)];
    float acc = a0 + a1;
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
    if (sub == 0) {
      const float sc = rowscale ? rowscale[r] : 1.0f;
      const float yh = acc * sc;
      const float yi = y[r];
      const float ev = yh > 0.0f ? yi / yh : 0.0f;
      e[r] = fold ? ev * sc : ev;
      if (yh > 0.0f) ll += (double)yi * log((double)yh) - (double)yh;
    }
  }
  const double t = block_sum_f64(ll, s_w);
  if (threadIdx.x == 0 && loglik) unsafeAtomicAdd(&loglik[*itp], t);
}

// Single frame, long rows: one wavefront per row with vector loads.
template <typename VT>
__global__ __launch_bounds__(256) void k_fwd_ratio_vec(int m, const int64_t* __restrict__ rowptr,
                                                       const int* __restrict__ col,
                                                       const typename VT::T* __restrict__ val,
                                                       const float* __restrict__ rowscale,
                                                       const float* __restrict__ x,
                                                       const float* __restrict__ y,
                                                       float* __restrict__ e, int fold,
                                                       const int* __restrict__ itp,
                                                       double* __restrict__ loglik) {
  __shared__ double s_w[256 / WAVE];
  const int lane = lane_id();
  const int wv = __builtin_amdgcn_readfirstlane(wave_id());
  const long long nw = (long long)gridDim.x * (256 / WAVE);
  double ll = 0.0;
  long long r = (long long)blockIdx.x * (256 / WAVE) + wv;
  long long pb = r < m ? row_bounds_load(rowptr, r, lane) : 0;
  for (; r < m; r += nw) {
    const long long b = lane_value_i64(pb, 0), end = lane_value_i64(pb, 1);
    const long long rn = r + nw;
    if (rn < m) pb = row_bounds_load(rowptr, rn, lane);
    const float acc = b < end ? wave_sum(row_dot_wave_s<VT, 16>(b, end, lane, col, val, x)) : 0.0f;
    if (lane == 0) {
      const float sc = rowscale ? rowscale[r] : 1.0f;
      const float yh = acc * sc;
      const float yi = y[r];
      const float ev = yh > 0.0f ? yi / yh : 0.0f;
      e[r] = fold ? ev * sc : ev;
      if (yh > 0.0f) ll += (double)yi * log((double)yh) - (double)yh;
    }
  }
  const double t = block_sum_f64(ll, s_w);
  if (threadIdx.x == 0 && loglik) unsafeAtomicAdd(&loglik[*itp], t);
}

// Multi-frame: one wavefront per row, lane = (slot, frame).
template <typename VT, int P>
__global__ __launch_bounds__(256) void k_fwd_ratio_mf(int m, int F,
                                                      const int64_t* __restrict__ rowptr,
                                                      const int* __restrict__ col,
                                                      const typename VT::T* __restrict__ val,
                                                      const float* __restrict__ rowscale,
                                                      const float* __restrict__ X,
                                                      const float* __restrict__ Y,
                                                      float* __restrict__ E,
                                                      const int* __restrict__ itp,
                                                      double* __restrict__ loglik) {
  constexpr int S = WAVE / P;
  const int lane = lane_id();
  const int f = lane % P, s = lane / P;
  const long long r = (long long)blockIdx.x * (256 / WAVE) + wave_id();
  if (r >= m) return;
  const long long b = rowptr[r], end = rowptr[r + 1];
  float acc = 0.0f;
  if (f < F)
    for (long long k = b + s; k < end; k += S)
      acc += VT::get(val, k) * X[(long long)__builtin_nontemporal_load(col + k) * F + f];
#pragma unroll
  for (int off = WAVE / 2; off >= P; off >>= 1) acc += __shfl_xor(acc, off, WAVE);
  if (s == 0 && f < F) {
    const float yh = acc * (rowscale ? rowscale[r] : 1.0f);
    const float yi = Y[r * F + f];
    E[r * F + f] = yh > 0.0f ? yi / yh : 0.0f;
    if (yh > 0.0f && loglik)
      unsafeAtomicAdd(&loglik[(long long)(*itp) * F + f], (double)yi * log((double)yh) - (double)yh);
  }
}

// x <- x c / (s + [s==0]) for active frames, FP64 sum per frame into slot it&1
// (grid-stride; per-frame partial sums in LDS, one global atomic per
// workgroup and frame)
__global__ __launch_bounds__(256) void k_update(float* __restrict__ x, const float* __restrict__ c,
                                                const float* __restrict__ s, long long n, int F,
                                                const int* __restrict__ itp,
                                                double* __restrict__ sum,
                                                const int* __restrict__ active) {
  __shared__ double s_sum[64];
  if (threadIdx.x < 64) s_sum[threadIdx.x] = 0.0;
  __syncthreads();
  const int slot = (*itp) & 1;
  const long long tot = n * F;
  double v1 = 0.0;   // single-frame running sum
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < tot;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long j = (F == 1) ? idx : idx / F;
    const int f = (F == 1) ? 0 : (int)(idx % F);
    float dvs = s[j];
    dvs += (float)(dvs == 0.0f);   // mlemOperations.hpp:48-50
    float xn = x[idx];
    if (!active || active[f]) { xn = xn * c[idx] / dvs; x[idx] = xn; }
    if (F == 1) v1 += xn;
    else atomicAdd(&s_sum[f], (double)xn);
  }
  if (F == 1) {
    v1 = wave_sum(v1);
    if (lane_id() == 0) atomicAdd(&s_sum[0], v1);
  }
  __syncthreads();
  if (threadIdx.x < F) unsafeAtomicAdd(&sum[slot * F + threadIdx.x], s_sum[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_normalise(float* __restrict__ x, float* __restrict__ c,
                                                   long long n, int F, const int* __restrict__ itp,
                                                   double* __restrict__ sum,
                                                   const int* __restrict__ active) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int slot = (*itp) & 1;
  if (idx < n * (long long)F) {
    const int f = (F == 1) ? 0 : (int)(idx % F);
    if (!active || active[f]) {
      const double S = sum[slot * F + f];
      x[idx] = x[idx] * (float)(1.0 / S);   // reco.cu:280-281
    }
    c[idx] = 0.0f;
  }
  if (idx < F) sum[(slot ^ 1) * F + idx] = 0.0;
}

__global__ void k_advance(int* itp) {
  if (threadIdx.x == 0) *itp += 1;
}

__global__ void k_fill_i32_one(int* p, int v) {
  if (threadIdx.x == 0) *p = v;
}

__global__ __launch_bounds__(256) void k_fill_f32(float* __restrict__ p, long long n, float v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_scale_by_sum(float* __restrict__ dst,
                                                      const float* __restrict__ src, long long n,
                                                      const double* __restrict__ sum) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i] * (float)(1.0 / (*sum));   // reco.cu:182-183
}

template <typename VT>
static int launch_fwd_ratio(hipStream_t st, int ncu, const epet_matrix* A, int F, const float* x,
                            const float* y, float* e, int fold, const int* itp, double* loglik) {
  using T = typename VT::T;
  const T* v = (const T*)A->val;
  const int m = A->m;
  if (m == 0) return EPET_OK;
  if (F == 1) {
    const int L = A->lanes_fwd > 0 ? A->lanes_fwd : pick_lanes(A->nnz, m);
    const long long cap = 16ll * ncu;
    if (L == 64 && A->lanes_fwd == 0) {
      k_fwd_ratio_vec<VT><<<(unsigned)std::min<long long>((m + 3) / 4, cap), 256, 0, st>>>(
          m, A->rowptr, A->col, v, A->rowscale, x, y, e, fold, itp, loglik);
      EPET_HIP(hipGetLastError());
      return EPET_OK;
    }
#define EPET_FR(LL)                                                                         \
  k_fwd_ratio<VT, LL><<<(unsigned)std::min<long long>(((long long)m + 256 / LL - 1) / (256 / LL), cap), \
                        256, 0, st>>>(m, A->rowptr, A->col, v, A->rowscale, x, y, e, fold, itp, loglik)
    switch (L) {
      case 64: EPET_FR(64); break;
      case 32: EPET_FR(32); break;
      case 16: EPET_FR(16); break;
      case 8: EPET_FR(8); break;
      default: EPET_FR(4); break;
    }
#undef EPET_FR
  } else {
    const unsigned g = grid_for(m, 4, 1ll << 31);
#define EPET_FM(P)                                                                            \
  k_fwd_ratio_mf<VT, P><<<g, 256, 0, st>>>(m, F, A->rowptr, A->col, v, A->rowscale, x, y, e, \
                                           itp, loglik)
    if (F <= 2) EPET_FM(2);
    else if (F <= 4) EPET_FM(4);
    else if (F <= 8) EPET_FM(8);
    else if (F <= 16) EPET_FM(16);
    else if (F <= 32) EPET_FM(32);
    else EPET_FM(64);
#undef EPET_FM
  }
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

// One MLEM iteration on `st` (capturable).  prepared_only: the single-frame
// back projection uses the plan A already holds and builds none (no
// allocation).
static int mlem_iteration(epet_ctx* ctx, const epet_matrix* A, int F, const float* y,
                          const float* s, float* x, float* e, float* c, int* itp, double* sum,
                          double* loglik, const int* active, bool prepared_only = false) {
  hipStream_t st = ctx->stream;
  int rc;
  if (F > 1 && A->tr) {
    const MfArgs a{F, x, y, e, c, itp, loglik};
    rc = use_blk(ctx, A, F) ? blk_iteration(ctx, A, a) : mf_iteration(ctx, A, a);
  } else {
    // single frame through the transposed copy: the ratio carries the row scale
    // and the back projection is a row-streaming product over the transpose
    const bool via_tr = F == 1 && bp_use_transpose(ctx, A);
    const int fold = via_tr && A->rowscale ? 1 : 0;
    if (F == 1 && lds_ready(ctx, A)) rc = lds_fwd_ratio(ctx, A, x, y, e, fold, itp, loglik);
    else switch (A->value_type) {
      case EPET_VAL_F32: rc = launch_fwd_ratio<ValF32>(st, ctx->num_cus, A, F, x, y, e, fold, itp, loglik); break;
      case EPET_VAL_BF16: rc = launch_fwd_ratio<ValBF16>(st, ctx->num_cus, A, F, x, y, e, fold, itp, loglik); break;
      default: rc = launch_fwd_ratio<ValF16>(st, ctx->num_cus, A, F, x, y, e, fold, itp, loglik); break;
    }
    if (rc != EPET_OK) return rc;
    if (via_tr) rc = matrix_bp_transpose(ctx, A->tr, 1.0f, e, 0.0f, c);
    else if (F == 1 && prepared_only) rc = matrix_spmv_t_prepared(ctx, A, 1.0f, e, 1.0f, c);
    else if (F == 1) rc = matrix_spmv(ctx, A, EPET_OP_T, 1.0f, e, 1.0f, c);
    else rc = matrix_spmm(ctx, A, EPET_OP_T, F, 1.0f, e, 1.0f, c);
  }
  if (rc != EPET_OK) return rc;
  const long long nv = (long long)A->n * F;
  k_update<<<grid_for(nv, 256, 4ll * ctx->num_cus), 256, 0, st>>>(x, c, s, A->n, F, itp, sum, active);
  k_normalise<<<grid_for(std::max<long long>(nv, F), 256, 1ll << 31), 256, 0, st>>>(x, c, A->n, F,
                                                                                   itp, sum, active);
  k_advance<<<1, 64, 0, st>>>(itp);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

// ------------------------------------------------------ allocation-free MLEM
// Working memory of epet_mlem_noalloc: one copy per GPU, held until the
// process ends.  In host code the array's name is only a stand-in; its device
// address on each GPU comes from hipGetSymbolAddress, taken once per device.
static __device__ __align__(256) unsigned char scratch[EPET_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on `device` (looked up on the first call for
// that device, with the device current).
static int scratch_address(int device, unsigned char** out) {
  EPET_CHECK(device >= 0 && device < NOALLOC_MAX_DEVICES, EPET_ERR_UNSUPPORTED,
             "device index beyond " + std::to_string(NOALLOC_MAX_DEVICES));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device] == nullptr) {
    int cur = -1;
    EPET_HIP(hipGetDevice(&cur));
    if (cur != device) EPET_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || p == nullptr)
      return fail(EPET_ERR_HIP, __FILE__, __LINE__,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[device];
  return EPET_OK;
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Byte offsets of the working arrays of epet_mlem_noalloc inside `scratch`:
// doubles {sum of s, x sums [2], log-likelihoods [niter]}, the iteration
// counter, e (m floats), c (n floats), normalised s (n floats).
struct NoallocLayout {
  size_t d, it, e, c, s, total;
};
static NoallocLayout noalloc_layout(long long m, long long n, int niter) {
  NoallocLayout L;
  L.d = 0;
  L.it = align256(sizeof(double) * (3 + (size_t)std::max(1, niter)));
  L.e = L.it + 256;
  L.c = align256(L.e + sizeof(float) * (size_t)std::max(1ll, m));
  L.s = align256(L.c + sizeof(float) * (size_t)std::max(1ll, n));
  L.total = align256(L.s + sizeof(float) * (size_t)std::max(1ll, n));
  return L;
}

}  // namespace epet

using namespace epet;

extern "C" {

int epet_divides(epet_ctx* ctx, float* res, const float* a, const float* b, long long n) {
  EPET_CHECK(ctx && (n == 0 || (res && a && b)), EPET_ERR_ARG, "null argument");
  if (n > 0) k_divides<<<grid_for(n, 256, 1ll << 31), 256, 0, ctx->stream>>>(res, a, b, n);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

int epet_dividesMultiplies(epet_ctx* ctx, float* res, const float* a, const float* b,
                           const float* c, long long n) {
  EPET_CHECK(ctx && (n == 0 || (res && a && b && c)), EPET_ERR_ARG, "null argument");
  if (n > 0)
    k_divides_multiplies<<<grid_for(n, 256, 1ll << 31), 256, 0, ctx->stream>>>(res, a, b, c, n);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

int epet_scales(epet_ctx* ctx, float* x, float alpha, long long n) {
  EPET_CHECK(ctx && (n == 0 || x), EPET_ERR_ARG, "null argument");
  if (n > 0) k_scales<<<grid_for(n, 256, 1ll << 31), 256, 0, ctx->stream>>>(x, alpha, n);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

int epet_sum(epet_ctx* ctx, const float* x, long long n, double* out) {
  EPET_CHECK(ctx && out && (n == 0 || x), EPET_ERR_ARG, "null argument");
  StreamTemps tmp(ctx->stream);
  double* d = nullptr;
  EPET_HIP(tmp.alloc(&d, sizeof(double)));
  k_zero_f64<<<1, 64, 0, ctx->stream>>>(d, 1);
  if (n > 0) k_sum_f64<<<grid_for(n, 256, 2048), 256, 0, ctx->stream>>>(x, n, d);
  EPET_HIP(hipGetLastError());
  return read_pinned(ctx, out, d, sizeof(double));
}

int epet_mlem(epet_ctx* ctx, const epet_matrix* A, int F, const float* y, const float* s_in,
              float* x, int niter, int use_graph, double tol, double* loglik_out,
              int* frame_iters) {
  EPET_CHECK(ctx && A && y && s_in && x, EPET_ERR_ARG, "null argument");
  EPET_CHECK(F >= 1 && F <= 64, EPET_ERR_UNSUPPORTED, "frames must be in [1, 64]");
  EPET_CHECK(niter >= 0, EPET_ERR_ARG, "negative iteration count");
  const long long nv = (long long)A->n * F;
  hipStream_t own = nullptr;
  hipStream_t saved = ctx->stream;
  if (use_graph && ctx->stream == nullptr) {
    // capture needs a non-null stream; order it after the caller's null-stream work
    EPET_HIP(hipStreamSynchronize(nullptr));
    EPET_HIP(hipStreamCreateWithFlags(&own, hipStreamNonBlocking));
    ctx->stream = own;
  }
  hipStream_t st = ctx->stream;
  // buffers: e (m F), c (n F), s (n), state
  float *e = nullptr, *c = nullptr, *s = nullptr;
  int* ibuf = nullptr;      // [it, active[F]]
  double* dbuf = nullptr;   // [ssum, sum[2F], loglik[niter F]]
  const size_t nd = 1 + 2 * (size_t)F + (size_t)std::max(1, niter) * F;
  int rc = EPET_OK;
  auto cleanup = [&]() {
    if (e) (void)hipFreeAsync(e, st);
    if (c) (void)hipFreeAsync(c, st);
    if (s) (void)hipFreeAsync(s, st);
    if (ibuf) (void)hipFreeAsync(ibuf, st);
    if (dbuf) (void)hipFreeAsync(dbuf, st);
    (void)hipStreamSynchronize(st);
    if (own) { (void)hipStreamDestroy(own); ctx->stream = saved; }
  };
#define EPET_MLEM_HIP(call)                                                              \
  do {                                                                                   \
    hipError_t _e = (call);                                                              \
    if (_e != hipSuccess) {                                                              \
      rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, std::string(#call) + ": " + hipGetErrorString(_e)); \
      cleanup();                                                                         \
      return rc;                                                                         \
    }                                                                                    \
  } while (0)
  EPET_MLEM_HIP(hipMallocAsync((void**)&e, sizeof(float) * std::max<long long>(1, (long long)A->m * F), st));
  EPET_MLEM_HIP(hipMallocAsync((void**)&c, sizeof(float) * std::max<long long>(1, nv), st));
  EPET_MLEM_HIP(hipMallocAsync((void**)&s, sizeof(float) * std::max(1, A->n), st));
  EPET_MLEM_HIP(hipMallocAsync((void**)&ibuf, sizeof(int) * (1 + F), st));
  EPET_MLEM_HIP(hipMallocAsync((void**)&dbuf, sizeof(double) * nd, st));
  int* itp = ibuf;
  int* active = ibuf + 1;
  double* ssum = dbuf;
  double* sum = dbuf + 1;
  double* ll = dbuf + 1 + 2 * F;
  k_zero_f64<<<grid_for((long long)nd, 256), 256, 0, st>>>(dbuf, (int)nd);
  k_fill_f32<<<grid_for(nv, 256, 1ll << 31), 256, 0, st>>>(c, nv, 0.0f);
  std::vector<int> hi(1 + F, 1);
  hi[0] = 0;
  EPET_MLEM_HIP(hipMemcpyAsync(ibuf, hi.data(), sizeof(int) * (1 + F), hipMemcpyHostToDevice, st));
  // s <- s / sum(s)  (reco.cu:182-183)
  k_sum_f64<<<grid_for(A->n, 256, 2048), 256, 0, st>>>(s_in, A->n, ssum);
  k_scale_by_sum<<<grid_for(A->n, 256, 1ll << 31), 256, 0, st>>>(s, s_in, A->n, ssum);
  EPET_MLEM_HIP(hipGetLastError());
  EPET_MLEM_HIP(hipStreamSynchronize(st));

  // back-projection plan (allocates) and multi-frame tuning before any capture
  if (F == 1) {
    rc = lds_prepare_map(ctx, const_cast<epet_matrix*>(A), true);
    if (rc == EPET_OK && bp_use_transpose(ctx, A)) rc = lds_prepare(ctx, A->tr);
    if (rc == EPET_OK && !bp_use_transpose(ctx, A)) rc = bp_prepare(ctx, const_cast<epet_matrix*>(A));
    if (rc != EPET_OK) { cleanup(); return rc; }
  } else if (F > 1 && A->tr && !ctx->opt.batched_gather) {
    // blocked copies (built once per padded frame count) and their launch shape
    rc = blk_prepare(ctx, const_cast<epet_matrix*>(A), F);
    // copies that could not be built leave the gather kernels in place
    if (rc == EPET_OK && blk_built(A, F) && !blk_ready(A, F)) {
      const MfArgs a{F, x, y, e, c, itp, nullptr};
      rc = blk_tune(ctx, const_cast<epet_matrix*>(A), a, 2);
    }
    if (rc != EPET_OK) { cleanup(); return rc; }
  }
  if (F > 1 && A->tr && !use_blk(ctx, A, F) && A->mf_F != F) {
    const MfArgs a{F, x, y, e, c, itp, nullptr};
    rc = mf_tune(ctx, const_cast<epet_matrix*>(A), a, 2);
    if (rc != EPET_OK) { cleanup(); return rc; }
  }
  k_fill_f32<<<grid_for(nv, 256, 1ll << 31), 256, 0, st>>>(c, nv, 0.0f);
  const bool stopping = tol > 0.0 && frame_iters != nullptr;
  std::vector<int> act(F, 1), iters(F, 0);
  std::vector<double> hll((size_t)std::max(1, niter) * F, 0.0);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (use_graph && niter > 0) {
    EPET_MLEM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = mlem_iteration(ctx, A, F, y, s, x, e, c, itp, sum, ll, active);
    hipError_t ce = hipStreamEndCapture(st, &graph);
    if (rc != EPET_OK) { cleanup(); return rc; }
    EPET_MLEM_HIP(ce);
    EPET_MLEM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  }
  for (int it = 0; it < niter; it++) {
    if (exec) {
      hipError_t ge = hipGraphLaunch(exec, st);
      if (ge != hipSuccess) { rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "hipGraphLaunch"); break; }
    } else {
      rc = mlem_iteration(ctx, A, F, y, s, x, e, c, itp, sum, ll, active);
      if (rc != EPET_OK) break;
    }
    if (stopping) {
      // per-frame stop on the relative log-likelihood change
      if (hipMemcpyAsync(hll.data() + (size_t)it * F, ll + (size_t)it * F, sizeof(double) * F,
                         hipMemcpyDeviceToHost, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess) {
        rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "log-likelihood readback");
        break;
      }
      bool changed = false, any = false;
      for (int f = 0; f < F; f++) {
        if (!act[f]) continue;
        iters[f] = it + 1;
        if (it > 0) {
          const double a = hll[(size_t)it * F + f], b = hll[(size_t)(it - 1) * F + f];
          if (std::fabs(a - b) <= tol * std::fabs(b)) { act[f] = 0; changed = true; }
        }
        any = any || act[f];
      }
      if (changed) {
        if (hipMemcpyAsync(active, act.data(), sizeof(int) * F, hipMemcpyHostToDevice, st) !=
            hipSuccess) {
          rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "frame flag upload");
          break;
        }
      }
      if (!any) break;
    }
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc == EPET_OK && loglik_out && niter > 0) {
    if (hipMemcpyAsync(loglik_out, ll, sizeof(double) * (size_t)niter * F, hipMemcpyDeviceToHost,
                       st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "log-likelihood readback");
  }
  if (frame_iters) {
    for (int f = 0; f < F; f++) frame_iters[f] = stopping ? iters[f] : niter;
  }
  cleanup();
#undef EPET_MLEM_HIP
  return rc;
}

int epet_mlem_noalloc(epet_ctx* ctx, const epet_matrix* A, const float* y, const float* s_in,
                      float* x, int niter, double* loglik_out) {
  EPET_CHECK(ctx && A && y && s_in && x, EPET_ERR_ARG, "null argument");
  EPET_CHECK(niter >= 0, EPET_ERR_ARG, "negative iteration count");
  const NoallocLayout L = noalloc_layout(A->m, A->n, niter);
  EPET_CHECK(L.total <= (size_t)EPET_NOALLOC_SCRATCH_BYTES, EPET_ERR_CAPACITY,
             "allocation-free MLEM needs " + std::to_string(L.total) +
                 " bytes of working memory (m + 2 n floats, niter doubles) for m = " +
                 std::to_string(A->m) + ", n = " + std::to_string(A->n) + ", niter = " +
                 std::to_string(niter) + "; the limit is " +
                 std::to_string((size_t)EPET_NOALLOC_SCRATCH_BYTES));
  unsigned char* base = nullptr;
  EPET_TRY(scratch_address(ctx->device, &base));
  hipStream_t st = ctx->stream;
  double* dbuf = (double*)(base + L.d);
  double* ssum = dbuf;
  double* sum = dbuf + 1;
  double* ll = dbuf + 3;
  int* itp = (int*)(base + L.it);
  float* e = (float*)(base + L.e);
  float* c = (float*)(base + L.c);
  float* s = (float*)(base + L.s);
  const int n = A->n;
  // the array keeps the previous call's contents: clear with kernels
  const int nd = 3 + std::max(1, niter);
  k_zero_f64<<<grid_for(nd, 256), 256, 0, st>>>(dbuf, nd);
  k_fill_i32_one<<<1, 64, 0, st>>>(itp, 0);
  k_fill_f32<<<grid_for(std::max(1, n), 256, 1ll << 31), 256, 0, st>>>(c, n, 0.0f);
  // s <- s / sum(s)  (reco.cu:182-183)
  if (n > 0) {
    k_sum_f64<<<grid_for(n, 256, 2048), 256, 0, st>>>(s_in, n, ssum);
    k_scale_by_sum<<<grid_for(n, 256, 1ll << 31), 256, 0, st>>>(s, s_in, n, ssum);
  }
  EPET_HIP(hipGetLastError());
  for (int it = 0; it < niter; it++)
    EPET_TRY(mlem_iteration(ctx, A, 1, y, s, x, e, c, itp, sum, ll, nullptr, true));
  if (loglik_out && niter > 0) {
    EPET_HIP(hipMemcpyAsync(loglik_out, ll, sizeof(double) * (size_t)niter, hipMemcpyDeviceToHost, st));
  }
  EPET_HIP(hipStreamSynchronize(st));
  return EPET_OK;
}

int epet_mlem_tune(epet_ctx* ctx, epet_matrix* A, int F, int* cfg, float* ms) {
  EPET_CHECK(ctx && A, EPET_ERR_ARG, "null argument");
  EPET_CHECK(F >= 2 && F <= 64, EPET_ERR_UNSUPPORTED, "frames must be in [2, 64]");
  EPET_CHECK(A->tr, EPET_ERR_STATE, "attach the transpose first (epet_matrix_transpose)");
  hipStream_t st = ctx->stream;
  float *x = nullptr, *y = nullptr, *e = nullptr, *c = nullptr;
  int* itp = nullptr;
  const long long nx = (long long)A->n * F, ny = (long long)A->m * F;
  auto release = [&]() {
    if (x) (void)hipFreeAsync(x, st);
    if (y) (void)hipFreeAsync(y, st);
    if (e) (void)hipFreeAsync(e, st);
    if (c) (void)hipFreeAsync(c, st);
    if (itp) (void)hipFreeAsync(itp, st);
  };
  if (hipMallocAsync((void**)&x, sizeof(float) * std::max(1ll, nx), st) != hipSuccess ||
      hipMallocAsync((void**)&c, sizeof(float) * std::max(1ll, nx), st) != hipSuccess ||
      hipMallocAsync((void**)&y, sizeof(float) * std::max(1ll, ny), st) != hipSuccess ||
      hipMallocAsync((void**)&e, sizeof(float) * std::max(1ll, ny), st) != hipSuccess ||
      hipMallocAsync((void**)&itp, sizeof(int), st) != hipSuccess) {
    (void)hipGetLastError();
    release();
    (void)hipStreamSynchronize(st);
    EPET_FAIL(EPET_ERR_ALLOC, "tuning buffers");
  }
  k_fill_f32<<<grid_for(nx, 256, 1ll << 31), 256, 0, st>>>(x, nx, 1.0f);
  k_fill_f32<<<grid_for(ny, 256, 1ll << 31), 256, 0, st>>>(y, ny, 1.0f);
  k_fill_i32_one<<<1, 64, 0, st>>>(itp, 0);
  const MfArgs a{F, x, y, e, c, itp, nullptr};
  int rc = EPET_OK;
  bool blocked = false;
  if (!ctx->opt.batched_gather) {
    rc = blk_prepare(ctx, A, F);
    if (rc == EPET_OK && blk_built(A, F)) {
      rc = blk_tune(ctx, A, a, 3);
      blocked = true;
    }
  }
  if (rc == EPET_OK && !blocked) {
    A->mf_F = 0;
    rc = mf_tune(ctx, A, a, 3);
  }
  release();
  EPET_HIP(hipStreamSynchronize(st));
  if (rc != EPET_OK) return rc;
  if (blocked) {
    int info[9];
    float t[2];
    EPET_TRY(blk_info(A, F, info, t));
    if (cfg) { cfg[0] = info[3]; cfg[1] = info[1]; cfg[2] = info[7]; cfg[3] = info[5]; }
    if (ms) { ms[0] = t[0]; ms[1] = t[1]; }
  } else {
    if (cfg) { cfg[0] = A->mf_fwd_R; cfg[1] = A->mf_fwd_P; cfg[2] = A->mf_bwd_R; cfg[3] = A->mf_bwd_P; }
    if (ms) { ms[0] = A->mf_fwd_ms; ms[1] = A->mf_bwd_ms; }
  }
  return EPET_OK;
}

int epet_mlem_kernels(epet_ctx* ctx, const epet_matrix* A, int F, int* info, float* ms) {
  EPET_CHECK(ctx && A && info, EPET_ERR_ARG, "null argument");
  for (int i = 0; i < 10; i++) info[i] = 0;
  float t[2] = {0.0f, 0.0f};
  if (F >= 2 && A->tr && use_blk(ctx, A, F)) {
    int bi[9];
    EPET_TRY(blk_info(A, F, bi, t));
    info[0] = 2;
    for (int i = 1; i < 9; i++) info[i] = bi[i];
  } else if (F >= 2 && A->tr && A->mf_F == F) {
    info[0] = 1;
    info[1] = A->mf_fwd_P; info[2] = A->mf_fwd_R; info[5] = A->mf_bwd_P; info[6] = A->mf_bwd_R;
    t[0] = A->mf_fwd_ms;
    t[1] = A->mf_bwd_ms;
  }
  if (ms) { ms[0] = t[0]; ms[1] = t[1]; }
  return EPET_OK;
}

}  // extern "C"
