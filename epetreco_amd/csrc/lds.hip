// lds.hip -- LDS-tiled row products y = A x (and, over a transposed copy, the
// back projection).
//
// The gathered vector x (n entries) is cut into tiles of LDS_TILE entries.
// Workgroup (tile t, row range g) of 1024 threads copies tile t into LDS, then
// its 16 wavefronts walk the rows of range g, R rows per wavefront at a time,
// summing each row's entries whose column lies in tile t (an entry range given
// by the per-matrix plan, since columns ascend within a row) with x read from
// LDS.  The row's partial for tile t goes to partial[t m + r]; a reduction sums
// the partials of every row in tile order (deterministic) and applies the row
// scale, alpha/beta or the MLEM ratio.  HBM sees the matrix stream once, x
// once per workgroup, and 8 bytes of plan and partial per (row, tile).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>

#include "epet_block.hpp"
#include "epet_internal.hpp"
#include "epet_rows.hpp"

namespace epet {

void launch_row_plan(hipStream_t st, int m, int T, int TS, const int64_t* rowptr, const int* col,
                     uint32_t* toff);

constexpr int LDS_BLOCK = 1024;
constexpr int LDS_WAVES = LDS_BLOCK / WAVE;

// R rows per wavefront (L = 64/R lanes per row), J entries per lane per pass;
// BWD = 1 marks the back projection over a transposed copy (a separate
// instantiation, so that profiles tell the two products apart); NT = 1 loads
// the entry stream with the non-temporal hint.  CI is the stored column type:
// int (global column) or uint16_t (tile-local column, col mod LDS_TILE).
template <typename VT, int R, int J, int BWD, int NT, typename CI>
__global__ __launch_bounds__(LDS_BLOCK) void k_rows_lds(int m, int n, int T, long long rows_per_wg,
                                                        const int64_t* __restrict__ rowptr,
                                                        const uint32_t* __restrict__ toff,
                                                        const CI* __restrict__ col,
                                                        const typename VT::T* __restrict__ val,
                                                        const float* __restrict__ x,
                                                        float* __restrict__ partial) {
  extern __shared__ float xs[];
  constexpr int L = WAVE / R;
  const int t = blockIdx.x % T;
  const long long g = blockIdx.x / T;
  const long long tb = (long long)t * LDS_TILE;
  const int cnt = (int)std::min<long long>(LDS_TILE, (long long)n - tb);
  // tile copy: 16-byte loads when the tile start is 16-byte aligned
  const int c4 = ((reinterpret_cast<uintptr_t>(x + tb) & 15) == 0) ? (cnt >> 2) : 0;
  for (int i = threadIdx.x; i < c4; i += LDS_BLOCK)
    reinterpret_cast<float4*>(xs)[i] = reinterpret_cast<const float4*>(x + tb)[i];
  for (int i = (c4 << 2) + threadIdx.x; i < cnt; i += LDS_BLOCK) xs[i] = x[tb + i];
  __syncthreads();
  const int lane = lane_id(), gi = lane / L, l = lane % L;
  const int w = __builtin_amdgcn_readfirstlane(wave_id());
  const long long r0 = g * rows_per_wg;
  const long long r1 = std::min<long long>((long long)m, r0 + rows_per_wg);
  const unsigned last = (unsigned)(cnt - 1);
  const long long step = (long long)LDS_WAVES * R;
  // segment [s0, s1) of row rbx + gi in tile t; empty past the range
  auto segment = [&](long long rbx, long long& s0, long long& s1) {
    s0 = 0;
    s1 = 0;
    const long long r = rbx + gi;
    if (r < r1) {
      const long long b = rowptr[r];
      const uint32_t* o = toff + r * (T + 1) + t;
      s0 = b + o[0];
      s1 = b + o[1];
    }
  };
  long long rb = r0 + (long long)w * R;
  long long n0, n1;
  segment(rb, n0, n1);
  for (; rb < r1; rb += step) {
    const long long r = rb + gi;
    const bool ok = r < r1;
    const long long s0 = n0, s1 = n1;
    segment(rb + step, n0, n1);   // the next segment's bounds, requested before this one's entries
    float acc = 0.0f;
    for (long long base = s0; base < s1; base += (long long)L * J) {
      CI c[J];
      float v[J];
      bool in[J];
#pragma unroll
      for (int j = 0; j < J; j++) {
        const long long k = base + l + (long long)j * L;
        in[j] = k < s1;
        const long long kc = in[j] ? k : s1 - 1;
        if (NT) {
          c[j] = __builtin_nontemporal_load(col + kc);
          v[j] = VT::get(val, kc);
        } else {
          c[j] = col[kc];
          v[j] = VT::get_cached(val, kc);
        }
      }
#pragma unroll
      for (int j = 0; j < J; j++) {
        // in-tile index; the clamp keeps LDS reads in range whatever the data
        const unsigned li =
            std::min(sizeof(CI) == 2 ? (unsigned)c[j] : (unsigned)((int)c[j] - (int)tb), last);
        acc += in[j] ? v[j] * xs[li] : 0.0f;
      }
    }
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, L);
    if (ok && l == 0) partial[(long long)t * m + r] = acc;
  }
}

__global__ __launch_bounds__(256) void k_col16(const int* __restrict__ col, long long nnz,
                                               uint16_t* __restrict__ col16) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
       k += (long long)gridDim.x * blockDim.x)
    col16[k] = (uint16_t)(col[k] & (LDS_TILE - 1));
}

// y_r = alpha * scale_r * sum_t partial[t][r] + beta * y_r
__global__ __launch_bounds__(256) void k_lds_reduce(int m, int T, const float* __restrict__ partial,
                                                    const float* __restrict__ rowscale,
                                                    float alpha, float beta,
                                                    float* __restrict__ y) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  float s = 0.0f;
  for (int t = 0; t < T; t++) s += partial[(long long)t * m + r];
  const float v = alpha * (rowscale ? rowscale[r] : 1.0f) * s;
  y[r] = beta == 0.0f ? v : v + beta * y[r];
}

// MLEM forward step from the partials: yhat_r = scale_r * sum_t partial[t][r],
// e_r = y_r / yhat_r (0 where yhat_r = 0; times the row scale when fold = 1),
// FP64 log-likelihood sum y log yhat - yhat, one atomic per workgroup.
__global__ __launch_bounds__(256) void k_lds_ratio(int m, int T, const float* __restrict__ partial,
                                                   const float* __restrict__ rowscale,
                                                   const float* __restrict__ y,
                                                   float* __restrict__ e, int fold,
                                                   const int* __restrict__ itp,
                                                   double* __restrict__ loglik) {
  __shared__ double s_w[256 / WAVE];
  double ll = 0.0;
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < m;
       r += (long long)gridDim.x * blockDim.x) {
    float s = 0.0f;
    for (int t = 0; t < T; t++) s += partial[(long long)t * m + r];
    const float sc = rowscale ? rowscale[r] : 1.0f;
    const float yh = s * sc;
    const float yi = y[r];
    const float ev = yh > 0.0f ? yi / yh : 0.0f;
    e[r] = fold ? ev * sc : ev;
    if (yh > 0.0f) ll += (double)yi * log((double)yh) - (double)yh;
  }
  ll = wave_sum(ll);
  if (lane_id() == 0) s_w[wave_id()] = ll;
  __syncthreads();
  if (threadIdx.x == 0 && loglik) {
    double tsum = 0.0;
    for (int w = 0; w < 256 / WAVE; w++) tsum += s_w[w];
    unsafeAtomicAdd(&loglik[*itp], tsum);
  }
}

template <typename VT, int R, int J, int BWD, int NT, typename CI>
static void set_lds_attr() {
  static bool done = false;
  if (!done) {
    (void)hipFuncSetAttribute((const void*)k_rows_lds<VT, R, J, BWD, NT, CI>,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              LDS_TILE * (int)sizeof(float));
    done = true;
  }
}

template <typename VT, int R, int J, int BWD, int NT, typename CI>
static void launch_rows_lds_c(epet_ctx* ctx, const epet_matrix* A, const float* x, const CI* col) {
  set_lds_attr<VT, R, J, BWD, NT, CI>();
  // each workgroup owns its CU's LDS; about four workgroups per CU in all, so
  // that tiles of unequal weight even out
  const long long G = std::max<long long>(1, (4ll * ctx->num_cus + A->lds_T - 1) / A->lds_T);
  const long long rows_per_wg = std::max<long long>(LDS_WAVES * R, (A->m + G - 1) / G);
  const long long ng = (A->m + rows_per_wg - 1) / rows_per_wg;
  k_rows_lds<VT, R, J, BWD, NT, CI><<<(unsigned)(ng * A->lds_T), LDS_BLOCK,
                                      LDS_TILE * sizeof(float), ctx->stream>>>(
      A->m, A->n, A->lds_T, rows_per_wg, A->rowptr, A->lds_toff, col,
      (const typename VT::T*)A->val, x, A->lds_partial);
}

// tile-local 16-bit columns when present (EPET_LDS_C16=0 keeps the int columns)
template <typename VT, int R, int J, int BWD, int NT>
static void launch_rows_lds(epet_ctx* ctx, const epet_matrix* A, const float* x) {
  const char* e = getenv("EPET_LDS_C16");
  if (A->lds_col16 && !(e && e[0] == '0'))
    launch_rows_lds_c<VT, R, J, BWD, NT, uint16_t>(ctx, A, x, A->lds_col16);
  else
    launch_rows_lds_c<VT, R, J, BWD, NT, int>(ctx, A, x, A->col);
}

// Rows per wavefront from the mean segment length.  Entries are loaded with
// ordinary (cached) loads: with R > 1 a row group's load covers only part of
// a cache line, and the line's other part is read by the next load, which a
// streaming hint would send back to HBM.  Measurement knobs: EPET_LDS_R (1, 2,
// 4 or 16) forces R, EPET_LDS_NT=1 selects non-temporal loads.
template <typename VT, int BWD, int NT>
static void rows_lds_r(epet_ctx* ctx, const epet_matrix* A, const float* x, int R) {
  if (R == 1) launch_rows_lds<VT, 1, 16, BWD, NT>(ctx, A, x);
  else if (R == 2) launch_rows_lds<VT, 2, 16, BWD, NT>(ctx, A, x);
  else if (R == 4) launch_rows_lds<VT, 4, 16, BWD, NT>(ctx, A, x);
  else launch_rows_lds<VT, 16, 16, BWD, NT>(ctx, A, x);
}

template <typename VT, int BWD>
static void rows_lds_dispatch(epet_ctx* ctx, const epet_matrix* A, const float* x) {
  const double seg = (double)A->nnz / std::max(1.0, (double)A->m * A->lds_T);
  int R = seg >= 384.0 ? 2 : seg >= 48.0 ? 4 : 16;
  const char* er = getenv("EPET_LDS_R");
  if (er) R = atoi(er);
  const char* en = getenv("EPET_LDS_NT");
  if (en && en[0] == '1') rows_lds_r<VT, BWD, 1>(ctx, A, x, R);
  else rows_lds_r<VT, BWD, 0>(ctx, A, x, R);
}

template <int BWD>
static void rows_lds(epet_ctx* ctx, const epet_matrix* A, const float* x) {
  switch (A->value_type) {
    case EPET_VAL_F32: rows_lds_dispatch<ValF32, BWD>(ctx, A, x); break;
    case EPET_VAL_BF16: rows_lds_dispatch<ValBF16, BWD>(ctx, A, x); break;
    default: rows_lds_dispatch<ValF16, BWD>(ctx, A, x); break;
  }
}

// Whether the LDS-tiled products apply (EPET_LDS=0 turns them off for
// measurements): long rows, ascending columns.
static bool lds_wanted(const epet_matrix* A) {
  const char* e = getenv("EPET_LDS");
  if (e && e[0] == '0') return false;
  return A->m > 0 && A->cols_strict && (double)A->nnz / A->m >= 64.0;
}

// Builds the plan and partial buffer once (allocates: call outside graph
// capture).  Leaves the matrix without a plan when allocation fails.
int lds_prepare(epet_ctx* ctx, epet_matrix* A) {
  if (A->lds_toff || A->lds_failed || !lds_wanted(A)) return EPET_OK;
  const int T = (int)(((long long)A->n + LDS_TILE - 1) / LDS_TILE);
  const long long nt = (long long)A->m * (T + 1);
  if (T < 1 || hipMalloc((void**)&A->lds_toff, sizeof(uint32_t) * nt) != hipSuccess ||
      hipMalloc((void**)&A->lds_partial, sizeof(float) * (size_t)A->m * T) != hipSuccess) {
    (void)hipGetLastError();
    if (A->lds_toff) (void)hipFree(A->lds_toff);
    A->lds_toff = nullptr;
    A->lds_partial = nullptr;
    A->lds_failed = true;
    return EPET_OK;
  }
  A->lds_T = T;
  launch_row_plan(ctx->stream, A->m, T, LDS_TILE, A->rowptr, A->col, A->lds_toff);
  // tile-local 16-bit columns (optional: without them the int columns are read)
  if (hipMalloc((void**)&A->lds_col16, sizeof(uint16_t) * (size_t)std::max(1ll, A->nnz)) ==
      hipSuccess) {
    if (A->nnz > 0)
      k_col16<<<grid_for(A->nnz, 256, 16ll * ctx->num_cus), 256, 0, ctx->stream>>>(A->col, A->nnz,
                                                                                   A->lds_col16);
  } else {
    (void)hipGetLastError();
    A->lds_col16 = nullptr;
  }
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

bool lds_ready(const epet_matrix* A) { return A->lds_toff != nullptr && lds_wanted(A); }

// y = alpha * diag(scale) A x + beta y (scale: the row scale when use_scale;
// bwd: A is a transposed copy)
int lds_spmv(epet_ctx* ctx, const epet_matrix* A, const float* x, float alpha, float beta,
             float* y, bool use_scale, bool bwd) {
  if (bwd) rows_lds<1>(ctx, A, x);
  else rows_lds<0>(ctx, A, x);
  k_lds_reduce<<<grid_for(A->m, 256, 1ll << 31), 256, 0, ctx->stream>>>(
      A->m, A->lds_T, A->lds_partial, use_scale ? A->rowscale : nullptr, alpha, beta, y);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

// MLEM forward step (capturable once prepared)
int lds_fwd_ratio(epet_ctx* ctx, const epet_matrix* A, const float* x, const float* y, float* e,
                  int fold, const int* itp, double* loglik) {
  rows_lds<0>(ctx, A, x);
  k_lds_ratio<<<grid_for(A->m, 256, 16ll * ctx->num_cus), 256, 0, ctx->stream>>>(
      A->m, A->lds_T, A->lds_partial, A->rowscale, y, e, fold, itp, loglik);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

}  // namespace epet
