// spmv.hip -- forward projection (A x), back projection (A^T e) and their
// multi-frame forms on ECSR matrices.
//
// Forward: a group of L lanes per row (L = 64/32/16/8/4), coalesced
// nontemporal streams of column ids and values, gathers of the image from L2,
// group reduction with cross-lane shuffles.  Back projection: the same row walk
// scattering a_ij * e_i into the image with float atomics in L2.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>

#include "epet_block.hpp"
#include "epet_internal.hpp"

namespace epet {

// stored value -> float (reduced-precision values are multiplied by the row
// scale by the caller)
struct ValF32 {
  using T = float;
  static __device__ __forceinline__ float get(const T* p, long long k) {
    return __builtin_nontemporal_load(p + k);
  }
};
struct ValBF16 {
  using T = uint16_t;
  static __device__ __forceinline__ float get(const T* p, long long k) {
    return __uint_as_float((uint32_t)__builtin_nontemporal_load(p + k) << 16);
  }
};
struct ValF16 {
  using T = uint16_t;
  static __device__ __forceinline__ float get(const T* p, long long k) {
    return __half2float(__ushort_as_half(__builtin_nontemporal_load(p + k)));
  }
};

template <typename IDX>
__device__ __forceinline__ int ldcol(const int* p, IDX k) {
  return __builtin_nontemporal_load(p + k);
}

// y[r] = alpha * scale_r * sum_k a_k x[col_k] + beta * y[r]
template <typename VT, typename RP, int L>
__global__ __launch_bounds__(256) void k_spmv_n(int m, const RP* __restrict__ rowptr,
                                                const int* __restrict__ col,
                                                const typename VT::T* __restrict__ val,
                                                const float* __restrict__ rowscale,
                                                const float* __restrict__ x, float alpha,
                                                float beta, float* __restrict__ y) {
  const int sub = threadIdx.x % L;
  const long long r = (long long)blockIdx.x * (256 / L) + threadIdx.x / L;
  if (r >= m) return;
  const long long b = rowptr[r], e = rowptr[r + 1];
  float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, acc3 = 0.0f;
  long long k = b + sub;
  for (; k + 3 * L < e; k += 4 * L) {
    const int c0 = ldcol(col, k), c1 = ldcol(col, k + L);
    const int c2 = ldcol(col, k + 2 * L), c3 = ldcol(col, k + 3 * L);
    const float v0 = VT::get(val, k), v1 = VT::get(val, k + L);
    const float v2 = VT::get(val, k + 2 * L), v3 = VT::get(val, k + 3 * L);
    acc0 += v0 * x[c0];
    acc1 += v1 * x[c1];
    acc2 += v2 * x[c2];
    acc3 += v3 * x[c3];
  }
  for (; k < e; k += L) acc0 += VT::get(val, k) * x[ldcol(col, k)];
  float acc = (acc0 + acc1) + (acc2 + acc3);
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
  if (sub == 0) {
    const float sc = rowscale ? rowscale[r] : 1.0f;
    const float v = alpha * sc * acc;
    y[r] = (beta == 0.0f) ? v : v + beta * y[r];
  }
}

// y[col_k] += alpha * scale_r * a_k * x[r]   (y pre-scaled by beta)
template <typename VT, typename RP, int L>
__global__ __launch_bounds__(256) void k_spmv_t(int m, const RP* __restrict__ rowptr,
                                                const int* __restrict__ col,
                                                const typename VT::T* __restrict__ val,
                                                const float* __restrict__ rowscale,
                                                const float* __restrict__ x, float alpha,
                                                float* __restrict__ y) {
  const int sub = threadIdx.x % L;
  const long long r = (long long)blockIdx.x * (256 / L) + threadIdx.x / L;
  if (r >= m) return;
  const float t = alpha * (rowscale ? rowscale[r] : 1.0f) * x[r];
  if (t == 0.0f) return;
  const long long b = rowptr[r], e = rowptr[r + 1];
  for (long long k = b + sub; k < e; k += L)
    unsafeAtomicAdd(&y[ldcol(col, k)], VT::get(val, k) * t);
}

__global__ __launch_bounds__(256) void k_scale_vec(float* __restrict__ y, long long n, float beta) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (beta == 0.0f) ? 0.0f : beta * y[i];
}

// Multi-frame forward: Y[r][f] = alpha * scale_r * sum_k a_k X[col_k][f] + beta Y.
// One wavefront per row; lane = (slot, frame) with P = frames rounded up to a
// power of two, 64/P entries of the row in flight per step.
template <typename VT, int P>
__global__ __launch_bounds__(256) void k_spmm_n(int m, int F, const int64_t* __restrict__ rowptr,
                                                const int* __restrict__ col,
                                                const typename VT::T* __restrict__ val,
                                                const float* __restrict__ rowscale,
                                                const float* __restrict__ X, float alpha,
                                                float beta, float* __restrict__ Y) {
  constexpr int S = WAVE / P;
  const int lane = lane_id();
  const int f = lane % P, s = lane / P;
  const long long r = (long long)blockIdx.x * (256 / WAVE) + wave_id();
  if (r >= m) return;
  const long long b = rowptr[r], e = rowptr[r + 1];
  float acc = 0.0f;
  if (f < F) {
    for (long long k = b + s; k < e; k += S) {
      const int c = col[k];
      acc += VT::get(val, k) * X[(long long)c * F + f];
    }
  }
#pragma unroll
  for (int off = WAVE / 2; off >= P; off >>= 1) acc += __shfl_xor(acc, off, WAVE);
  if (s == 0 && f < F) {
    const float sc = rowscale ? rowscale[r] : 1.0f;
    const float v = alpha * sc * acc;
    const long long o = r * F + f;
    Y[o] = (beta == 0.0f) ? v : v + beta * Y[o];
  }
}

// Multi-frame back projection: Y[col_k][f] += alpha * scale_r * a_k * X[r][f].
template <typename VT, int P>
__global__ __launch_bounds__(256) void k_spmm_t(int m, int F, const int64_t* __restrict__ rowptr,
                                                const int* __restrict__ col,
                                                const typename VT::T* __restrict__ val,
                                                const float* __restrict__ rowscale,
                                                const float* __restrict__ X, float alpha,
                                                float* __restrict__ Y) {
  constexpr int S = WAVE / P;
  const int lane = lane_id();
  const int f = lane % P, s = lane / P;
  const long long r = (long long)blockIdx.x * (256 / WAVE) + wave_id();
  if (r >= m) return;
  if (f >= F) return;
  const float t = alpha * (rowscale ? rowscale[r] : 1.0f) * X[r * F + f];
  const long long b = rowptr[r], e = rowptr[r + 1];
  for (long long k = b + s; k < e; k += S) {
    const int c = col[k];
    unsafeAtomicAdd(&Y[(long long)c * F + f], VT::get(val, k) * t);
  }
}

// lanes per row from the average row length
int pick_lanes(long long nnz, int m) {
  const double avg = m > 0 ? (double)nnz / m : 0.0;
  if (avg >= 96) return 64;
  if (avg >= 48) return 32;
  if (avg >= 24) return 16;
  if (avg >= 12) return 8;
  return 4;
}

template <typename VT, typename RP>
static int launch_spmv(hipStream_t st, int op, int m, int n, long long nnz, const RP* rowptr,
                       const int* col, const void* val, const float* rowscale, const float* x,
                       float alpha, float beta, float* y, int lanes) {
  using T = typename VT::T;
  const T* v = (const T*)val;
  if (lanes <= 0) lanes = pick_lanes(nnz, m);
  if (op == EPET_OP_N) {
    if (m == 0) return EPET_OK;
#define EPET_LN(L)                                                                       \
  k_spmv_n<VT, RP, L><<<grid_for(m, 256 / L, 1ll << 31), 256, 0, st>>>(m, rowptr, col, v, \
                                                                      rowscale, x, alpha, \
                                                                      beta, y)
    switch (lanes) {
      case 64: EPET_LN(64); break;
      case 32: EPET_LN(32); break;
      case 16: EPET_LN(16); break;
      case 8: EPET_LN(8); break;
      default: EPET_LN(4); break;
    }
#undef EPET_LN
  } else {
    if (beta != 1.0f && n > 0)
      k_scale_vec<<<grid_for(n, 256, 1ll << 31), 256, 0, st>>>(y, n, beta);
    if (m == 0) return EPET_OK;
#define EPET_LT(L)                                                                       \
  k_spmv_t<VT, RP, L><<<grid_for(m, 256 / L, 1ll << 31), 256, 0, st>>>(m, rowptr, col, v, \
                                                                      rowscale, x, alpha, y)
    switch (lanes) {
      case 64: EPET_LT(64); break;
      case 32: EPET_LT(32); break;
      case 16: EPET_LT(16); break;
      case 8: EPET_LT(8); break;
      default: EPET_LT(4); break;
    }
#undef EPET_LT
  }
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

int matrix_spmv(epet_ctx* ctx, const epet_matrix* A, int op, float alpha, const float* x,
                float beta, float* y) {
  switch (A->value_type) {
    case EPET_VAL_F32:
      return launch_spmv<ValF32, int64_t>(ctx->stream, op, A->m, A->n, A->nnz, A->rowptr, A->col,
                                          A->val, nullptr, x, alpha, beta, y, A->lanes_fwd);
    case EPET_VAL_BF16:
      return launch_spmv<ValBF16, int64_t>(ctx->stream, op, A->m, A->n, A->nnz, A->rowptr, A->col,
                                           A->val, A->rowscale, x, alpha, beta, y, A->lanes_fwd);
    case EPET_VAL_F16:
      return launch_spmv<ValF16, int64_t>(ctx->stream, op, A->m, A->n, A->nnz, A->rowptr, A->col,
                                          A->val, A->rowscale, x, alpha, beta, y, A->lanes_fwd);
  }
  EPET_FAIL(EPET_ERR_STATE, "unknown value type");
}

template <typename VT>
static int launch_spmm(hipStream_t st, const epet_matrix* A, int op, int F, float alpha,
                       const float* X, float beta, float* Y) {
  using T = typename VT::T;
  const T* v = (const T*)A->val;
  const int m = A->m;
  if (op == EPET_OP_T && beta != 1.0f)
    k_scale_vec<<<grid_for((long long)A->n * F, 256, 1ll << 31), 256, 0, st>>>(Y, (long long)A->n * F,
                                                                            beta);
  if (m == 0) return EPET_OK;
  const unsigned g = grid_for(m, 4, 1ll << 31);
#define EPET_SP(P)                                                                          \
  do {                                                                                      \
    if (op == EPET_OP_N)                                                                    \
      k_spmm_n<VT, P><<<g, 256, 0, st>>>(m, F, A->rowptr, A->col, v, A->rowscale, X, alpha, \
                                         beta, Y);                                          \
    else                                                                                    \
      k_spmm_t<VT, P><<<g, 256, 0, st>>>(m, F, A->rowptr, A->col, v, A->rowscale, X, alpha, \
                                         Y);                                                \
  } while (0)
  if (F <= 1) EPET_SP(1);
  else if (F <= 2) EPET_SP(2);
  else if (F <= 4) EPET_SP(4);
  else if (F <= 8) EPET_SP(8);
  else if (F <= 16) EPET_SP(16);
  else if (F <= 32) EPET_SP(32);
  else EPET_SP(64);
#undef EPET_SP
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

int matrix_spmm(epet_ctx* ctx, const epet_matrix* A, int op, int F, float alpha, const float* X,
                float beta, float* Y) {
  switch (A->value_type) {
    case EPET_VAL_F32: return launch_spmm<ValF32>(ctx->stream, A, op, F, alpha, X, beta, Y);
    case EPET_VAL_BF16: return launch_spmm<ValBF16>(ctx->stream, A, op, F, alpha, X, beta, Y);
    case EPET_VAL_F16: return launch_spmm<ValF16>(ctx->stream, A, op, F, alpha, X, beta, Y);
  }
  EPET_FAIL(EPET_ERR_STATE, "unknown value type");
}

}  // namespace epet

using namespace epet;

extern "C" {

int epet_csrmv(epet_ctx* ctx, int op, int m, int n, long long nnz, const float* alpha,
               const float* val, const int* rowptr, const int* col, const float* x,
               const float* beta, float* y) {
  EPET_CHECK(ctx && alpha && beta && rowptr && x && y && (nnz == 0 || (val && col)),
             EPET_ERR_ARG, "null argument");
  EPET_CHECK(op == EPET_OP_N || op == EPET_OP_T, EPET_ERR_ARG, "unknown operation");
  EPET_CHECK(m >= 0 && n >= 0 && nnz >= 0, EPET_ERR_ARG, "bad sizes");
  return launch_spmv<ValF32, int>(ctx->stream, op, m, n, nnz, rowptr, col, val, nullptr, x,
                                  *alpha, *beta, y, 0);
}

int epet_matrix_spmv(epet_ctx* ctx, const epet_matrix* A, int op, float alpha, const float* x,
                     float beta, float* y) {
  EPET_CHECK(ctx && A && x && y, EPET_ERR_ARG, "null argument");
  EPET_CHECK(op == EPET_OP_N || op == EPET_OP_T, EPET_ERR_ARG, "unknown operation");
  return matrix_spmv(ctx, A, op, alpha, x, beta, y);
}

int epet_matrix_spmm(epet_ctx* ctx, const epet_matrix* A, int op, int nrhs, float alpha,
                     const float* X, float beta, float* Y) {
  EPET_CHECK(ctx && A && X && Y, EPET_ERR_ARG, "null argument");
  EPET_CHECK(op == EPET_OP_N || op == EPET_OP_T, EPET_ERR_ARG, "unknown operation");
  EPET_CHECK(nrhs >= 1 && nrhs <= 64, EPET_ERR_UNSUPPORTED, "nrhs must be in [1, 64]");
  return matrix_spmm(ctx, A, op, nrhs, alpha, X, beta, Y);
}

}  // extern "C"
