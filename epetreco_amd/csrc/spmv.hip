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
#include <cstdlib>
#include <string>

#include "epet_block.hpp"
#include "epet_internal.hpp"
#include "epet_rows.hpp"

namespace epet {

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

// Forward projection, one wavefront per row, persistent grid-stride over rows.
// Row indices are wave-uniform; the next row's bounds are requested before the
// current row is summed.  BWD = 1 is the same product over the transposed copy
// (the back projection), instantiated apart so that profiles tell them apart.
template <typename VT, int U, int BWD>
__global__ __launch_bounds__(256) void k_spmv_n_vec(int m, const int64_t* __restrict__ rowptr,
                                                    const int* __restrict__ col,
                                                    const typename VT::T* __restrict__ val,
                                                    const float* __restrict__ rowscale,
                                                    const float* __restrict__ x, float alpha,
                                                    float beta, float* __restrict__ y) {
  const int lane = lane_id();
  const int wv = __builtin_amdgcn_readfirstlane(wave_id());
  const long long nw = (long long)gridDim.x * (256 / WAVE);
  long long r = (long long)blockIdx.x * (256 / WAVE) + wv;
  long long pb = r < m ? row_bounds_load(rowptr, r, lane) : 0;
  for (; r < m; r += nw) {
    const long long b = lane_value_i64(pb, 0), e = lane_value_i64(pb, 1);
    const long long rn = r + nw;
    if (rn < m) pb = row_bounds_load(rowptr, rn, lane);
    float acc = 0.0f;
    if (b < e) acc = wave_sum(row_dot_wave_s<VT, U>(b, e, lane, col, val, x));
    if (lane == 0) {
      const float v = alpha * (rowscale ? rowscale[r] : 1.0f) * acc;
      y[r] = (beta == 0.0f) ? v : v + beta * y[r];
    }
  }
}

// ------------------------------------------------------------------ tiled back projection
// The image is cut into tiles of BP_TILE consecutive voxel ids (64 KB of LDS).
// Workgroup (row block, tile) accumulates a_ij e_i of its rows' entries inside
// the tile with LDS float atomics and writes the tile into its row block's
// partial image; a reduction sums the partial images.  Entry ranges of every
// (row, tile) come from a per-matrix plan of tile-boundary offsets.
constexpr int BP_TILE = 16384;
constexpr int BP_BLOCK = 1024;

__global__ __launch_bounds__(256) void k_bp_plan(int m, int T, int TS,
                                                 const int64_t* __restrict__ rowptr,
                                                 const int* __restrict__ col,
                                                 uint32_t* __restrict__ toff) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)m * (T + 1)) return;
  const long long r = idx / (T + 1);
  const int t = (int)(idx % (T + 1));
  const long long b = rowptr[r], e = rowptr[r + 1];
  long long lo = b, hi = e;   // first entry with col >= t*TS
  const long long target = (long long)t * TS;
  if (t == T) lo = e;
  else
    while (lo < hi) {
      const long long mid = (lo + hi) >> 1;
      if ((long long)col[mid] < target) lo = mid + 1; else hi = mid;
    }
  toff[idx] = (uint32_t)(lo - b);
}

template <typename VT>
__global__ __launch_bounds__(BP_BLOCK) void k_bp_tiled(int m, int n, int T, int rows_per_block,
                                                       const int64_t* __restrict__ rowptr,
                                                       const int* __restrict__ col,
                                                       const typename VT::T* __restrict__ val,
                                                       const float* __restrict__ rowscale,
                                                       const uint32_t* __restrict__ toff,
                                                       const float* __restrict__ e, float alpha,
                                                       float* __restrict__ partial) {
  __shared__ float acc[BP_TILE];
  const int t = blockIdx.y;
  const long long r0 = (long long)blockIdx.x * rows_per_block;
  const long long r1 = min((long long)m, r0 + rows_per_block);
  const int tile0 = t * BP_TILE;
  const int tn = min(BP_TILE, n - tile0);
  for (int j = threadIdx.x; j < BP_TILE; j += BP_BLOCK) acc[j] = 0.0f;
  __syncthreads();
  const int sub = threadIdx.x & 31;          // 32 lanes per row segment
  const int seg = threadIdx.x >> 5;          // 32 segments per workgroup
  for (long long r = r0 + seg; r < r1; r += BP_BLOCK / 32) {
    float ev = e[r];
    if (ev == 0.0f) continue;
    ev *= alpha * (rowscale ? rowscale[r] : 1.0f);
    const long long base = rowptr[r];
    const long long lo = base + toff[r * (T + 1) + t];
    const long long hi = base + toff[r * (T + 1) + t + 1];
    long long k = lo + sub;
    for (; k + 96 < hi; k += 128) {
      const int c0 = ldcol(col, k), c1 = ldcol(col, k + 32);
      const int c2 = ldcol(col, k + 64), c3 = ldcol(col, k + 96);
      const float v0 = VT::get(val, k), v1 = VT::get(val, k + 32);
      const float v2 = VT::get(val, k + 64), v3 = VT::get(val, k + 96);
      // the plan puts every column of [lo, hi) inside the tile; the unsigned
      // compare keeps LDS indices in range even for an inconsistent plan
      const unsigned j0 = c0 - tile0, j1 = c1 - tile0, j2 = c2 - tile0, j3 = c3 - tile0;
      if (j0 < BP_TILE) atomicAdd(&acc[j0], v0 * ev);
      if (j1 < BP_TILE) atomicAdd(&acc[j1], v1 * ev);
      if (j2 < BP_TILE) atomicAdd(&acc[j2], v2 * ev);
      if (j3 < BP_TILE) atomicAdd(&acc[j3], v3 * ev);
    }
    for (; k < hi; k += 32) {
      const unsigned j = ldcol(col, k) - tile0;
      if (j < BP_TILE) atomicAdd(&acc[j], VT::get(val, k) * ev);
    }
  }
  __syncthreads();
  float* out = partial + (size_t)blockIdx.x * n + tile0;
  for (int j = threadIdx.x; j < tn; j += BP_BLOCK) out[j] = acc[j];
}

// y[j] = beta * y[j] + sum over row blocks of partial[b][j]
__global__ __launch_bounds__(256) void k_bp_reduce(int n, int nblk,
                                                   const float* __restrict__ partial,
                                                   float beta, float* __restrict__ y) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
  int b = 0;
  for (; b + 3 < nblk; b += 4) {
    s0 += partial[(size_t)b * n + j];
    s1 += partial[(size_t)(b + 1) * n + j];
    s2 += partial[(size_t)(b + 2) * n + j];
    s3 += partial[(size_t)(b + 3) * n + j];
  }
  for (; b < nblk; b++) s0 += partial[(size_t)b * n + j];
  const float s = (s0 + s1) + (s2 + s3);
  y[j] = (beta == 0.0f) ? s : beta * y[j] + s;
}

int bp_prepare(epet_ctx* ctx, epet_matrix* A) {
  if (A->bp_toff || A->bp_failed) return EPET_OK;
  if (A->m == 0 || A->n == 0) { A->bp_failed = true; return EPET_OK; }
  const int T = (A->n + BP_TILE - 1) / BP_TILE;
  const int target_wgs = 8 * ctx->num_cus;
  int nblk = std::max(1, std::min(A->m, (target_wgs + T - 1) / T));
  const size_t toff_bytes = sizeof(uint32_t) * (size_t)A->m * (T + 1);
  const size_t part_bytes = sizeof(float) * (size_t)nblk * A->n;
  if (hipMalloc((void**)&A->bp_toff, toff_bytes) != hipSuccess ||
      hipMalloc((void**)&A->bp_partial, part_bytes) != hipSuccess) {
    (void)hipGetLastError();
    if (A->bp_toff) (void)hipFree(A->bp_toff);
    A->bp_toff = nullptr;
    A->bp_partial = nullptr;
    A->bp_failed = true;   // fall back to the atomic kernel
    return EPET_OK;
  }
  A->bp_T = T;
  A->bp_tile = BP_TILE;
  A->bp_nblk = nblk;
  const long long nthreads = (long long)A->m * (T + 1);
  k_bp_plan<<<grid_for(nthreads, 256, 1ll << 31), 256, 0, ctx->stream>>>(A->m, T, BP_TILE, A->rowptr,
                                                                        A->col, A->bp_toff);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

template <typename VT>
static int launch_bp_tiled(hipStream_t st, const epet_matrix* A, const float* x, float alpha,
                           float beta, float* y) {
  using T = typename VT::T;
  const int rows_per_block = (A->m + A->bp_nblk - 1) / A->bp_nblk;
  dim3 grid(A->bp_nblk, A->bp_T);
  k_bp_tiled<VT><<<grid, BP_BLOCK, 0, st>>>(A->m, A->n, A->bp_T, rows_per_block, A->rowptr, A->col,
                                            (const T*)A->val, A->rowscale, A->bp_toff, x, alpha,
                                            A->bp_partial);
  k_bp_reduce<<<grid_for(A->n, 256, 1ll << 31), 256, 0, st>>>(A->n, A->bp_nblk, A->bp_partial,
                                                             beta, y);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

template <typename VT, int BWD>
static int launch_fwd_vec(epet_ctx* ctx, const epet_matrix* A, const float* x, float alpha,
                          float beta, float* y) {
  using T = typename VT::T;
  // 16 workgroups per CU, persistent grid-stride over rows; lane-strided
  // layout with 16 entries per lane per pass
  const long long want = ((long long)A->m + 3) / 4;
  const unsigned g = (unsigned)std::max<long long>(1, std::min<long long>(want, 16ll * ctx->num_cus));
  k_spmv_n_vec<VT, 16, BWD><<<g, 256, 0, ctx->stream>>>(A->m, A->rowptr, A->col, (const T*)A->val,
                                                        A->rowscale, x, alpha, beta, y);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
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

// Back-projection choice: over the attached transpose (LDS-tiled, blocked or
// row kernel) unless the context asks for the atomic kernel or the LDS image
// tiles over A itself.
bool bp_use_transpose(const epet_ctx* ctx, const epet_matrix* A) {
  return A->tr && ctx->opt.bp != BP_ATOMIC && ctx->opt.bp != BP_TILED;
}

// Blocked back projection over a transposed copy: the columns of At (channel
// rows of A) are cut into XB_BLOCKS contiguous blocks and workgroup b sums the
// entries of block b mod XB_BLOCKS of its rows.  Workgroups are dispatched
// round-robin over the 8 XCDs, so each XCD gathers only its block of x (the
// ratio vector), which stays resident in its L2; the placement only affects
// speed.  Partial products are summed in block order by k_bp_xb_reduce.
template <typename VT, int J>
__global__ __launch_bounds__(256) void k_bp_xb(int n, const int64_t* __restrict__ tptr,
                                               const uint32_t* __restrict__ toff,
                                               const int* __restrict__ tcol,
                                               const typename VT::T* __restrict__ tval,
                                               const float* __restrict__ x,
                                               float* __restrict__ partial) {
  const int lane = lane_id();
  const int kb = blockIdx.x % XB_BLOCKS;
  const int wv = __builtin_amdgcn_readfirstlane(wave_id());
  const long long nw = (long long)(gridDim.x / XB_BLOCKS) * (256 / WAVE);
  for (long long j = (long long)(blockIdx.x / XB_BLOCKS) * (256 / WAVE) + wv; j < n; j += nw) {
    const long long rb = tptr[j];
    const uint32_t* o = toff + j * (XB_BLOCKS + 1) + kb;
    const long long b = rb + o[0], e = rb + o[1];
    float acc = 0.0f;
    if (b < e) acc = wave_sum(row_dot_wave_s<VT, J>(b, e, lane, tcol, tval, x));
    if (lane == 0) partial[(long long)kb * n + j] = acc;
  }
}

__global__ __launch_bounds__(256) void k_bp_xb_reduce(int n, const float* __restrict__ partial,
                                                      float alpha, float beta,
                                                      float* __restrict__ y) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float s = 0.0f;
#pragma unroll
  for (int k = 0; k < XB_BLOCKS; k++) s += partial[(long long)k * n + j];
  y[j] = beta == 0.0f ? alpha * s : alpha * s + beta * y[j];
}

// Entry offsets of every column-tile boundary (tiles of TS columns) in every
// row: toff[r (T+1) + t] = first entry of row r with column >= t TS.
void launch_row_plan(hipStream_t st, int m, int T, int TS, const int64_t* rowptr, const int* col,
                     uint32_t* toff) {
  const long long nt = (long long)m * (T + 1);
  if (nt > 0) k_bp_plan<<<grid_for(nt, 256, 1ll << 31), 256, 0, st>>>(m, T, TS, rowptr, col, toff);
}

// Plan of the blocked back projection for a transposed copy (entry offsets of
// every column-block boundary in every row); left absent when it cannot be
// allocated.
void xb_prepare(epet_ctx* ctx, epet_matrix* At) {
  if (At->xb_toff || At->m == 0) return;
  const long long nt = (long long)At->m * (XB_BLOCKS + 1);
  if (hipMalloc((void**)&At->xb_toff, sizeof(uint32_t) * nt) != hipSuccess) {
    (void)hipGetLastError();
    At->xb_toff = nullptr;
    return;
  }
  if (hipMalloc((void**)&At->xb_partial, sizeof(float) * (size_t)At->m * XB_BLOCKS) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(At->xb_toff);
    At->xb_toff = nullptr;
    At->xb_partial = nullptr;
    return;
  }
  const int bs = (At->n + XB_BLOCKS - 1) / XB_BLOCKS;
  k_bp_plan<<<grid_for(nt, 256, 1ll << 31), 256, 0, ctx->stream>>>(At->m, XB_BLOCKS, bs, At->rowptr,
                                                                   At->col, At->xb_toff);
}

int lds_prepare(epet_ctx* ctx, epet_matrix* A);
int lds_prepare_map(epet_ctx* ctx, epet_matrix* A, bool voxel_columns);
bool lds_ready(const epet_ctx* ctx, const epet_matrix* A);
int lds_spmv(epet_ctx* ctx, const epet_matrix* A, const float* x, float alpha, float beta,
             float* y, bool use_scale, bool bwd);

// y = alpha At x + beta y over the rows of a transposed copy At (unscaled values)
template <typename VT>
static int bp_rows(epet_ctx* ctx, const epet_matrix* At, float alpha, const float* x, float beta,
                   float* y) {
  if (lds_ready(ctx, At)) return lds_spmv(ctx, At, x, alpha, beta, y, false, true);
  const double avg = At->m ? (double)At->nnz / At->m : 0.0;
  if (At->xb_toff && avg >= 512.0 && ctx->opt.bp != BP_ROWS) {
    const unsigned g = (unsigned)(XB_BLOCKS * std::max(1, 2 * ctx->num_cus));
    k_bp_xb<VT, 16><<<g, 256, 0, ctx->stream>>>(At->m, At->rowptr, At->xb_toff, At->col,
                                                 (const typename VT::T*)At->val, x, At->xb_partial);
    k_bp_xb_reduce<<<grid_for(At->m, 256, 1ll << 31), 256, 0, ctx->stream>>>(At->m, At->xb_partial,
                                                                            alpha, beta, y);
    EPET_HIP(hipGetLastError());
    return EPET_OK;
  }
  if (avg >= 64.0) return launch_fwd_vec<VT, 1>(ctx, At, x, alpha, beta, y);
  return launch_spmv<VT, int64_t>(ctx->stream, EPET_OP_N, At->m, At->n, At->nnz, At->rowptr,
                                  At->col, At->val, nullptr, x, alpha, beta, y, 0);
}

template <typename VT>
static int spmv_dispatch(epet_ctx* ctx, const epet_matrix* A, int op, float alpha, const float* x,
                         float beta, float* y) {
  const double avg = A->m ? (double)A->nnz / A->m : 0.0;
  if (op == EPET_OP_N) {
    EPET_TRY(lds_prepare_map(ctx, const_cast<epet_matrix*>(A), true));
    if (lds_ready(ctx, A)) return lds_spmv(ctx, A, x, alpha, beta, y, true, false);
  }
  if (op == EPET_OP_N && avg >= 64.0 && A->lanes_fwd == 0)
    return launch_fwd_vec<VT, 0>(ctx, A, x, alpha, beta, y);
  if (op == EPET_OP_T && !A->rowscale && bp_use_transpose(ctx, A)) {
    EPET_TRY(lds_prepare(ctx, A->tr));
    return bp_rows<VT>(ctx, A->tr, alpha, x, beta, y);
  }
  if (op == EPET_OP_T && ctx->opt.bp != BP_ATOMIC) {
    EPET_TRY(bp_prepare(ctx, const_cast<epet_matrix*>(A)));
    if (A->bp_toff) return launch_bp_tiled<VT>(ctx->stream, A, x, alpha, beta, y);
  }
  return launch_spmv<VT, int64_t>(ctx->stream, op, A->m, A->n, A->nnz, A->rowptr, A->col, A->val,
                                  A->rowscale, x, alpha, beta, y, A->lanes_fwd);
}

int matrix_spmv(epet_ctx* ctx, const epet_matrix* A, int op, float alpha, const float* x,
                float beta, float* y) {
  switch (A->value_type) {
    case EPET_VAL_F32: return spmv_dispatch<ValF32>(ctx, A, op, alpha, x, beta, y);
    case EPET_VAL_BF16: return spmv_dispatch<ValBF16>(ctx, A, op, alpha, x, beta, y);
    case EPET_VAL_F16: return spmv_dispatch<ValF16>(ctx, A, op, alpha, x, beta, y);
  }
  EPET_FAIL(EPET_ERR_STATE, "unknown value type");
}

// y = alpha A^T x + beta y with the back-projection plan A already holds (the
// transposed copy's, the tiled plan, or none: the atomic kernel).  Builds no
// plan, so it allocates nothing.
template <typename VT>
static int spmv_t_prepared(epet_ctx* ctx, const epet_matrix* A, float alpha, const float* x,
                           float beta, float* y) {
  if (!A->rowscale && bp_use_transpose(ctx, A)) return bp_rows<VT>(ctx, A->tr, alpha, x, beta, y);
  if (ctx->opt.bp != BP_ATOMIC && A->bp_toff) return launch_bp_tiled<VT>(ctx->stream, A, x, alpha, beta, y);
  return launch_spmv<VT, int64_t>(ctx->stream, EPET_OP_T, A->m, A->n, A->nnz, A->rowptr, A->col,
                                  A->val, A->rowscale, x, alpha, beta, y, A->lanes_fwd);
}

int matrix_spmv_t_prepared(epet_ctx* ctx, const epet_matrix* A, float alpha, const float* x,
                           float beta, float* y) {
  switch (A->value_type) {
    case EPET_VAL_F32: return spmv_t_prepared<ValF32>(ctx, A, alpha, x, beta, y);
    case EPET_VAL_BF16: return spmv_t_prepared<ValBF16>(ctx, A, alpha, x, beta, y);
    case EPET_VAL_F16: return spmv_t_prepared<ValF16>(ctx, A, alpha, x, beta, y);
  }
  EPET_FAIL(EPET_ERR_STATE, "unknown value type");
}

int matrix_bp_transpose(epet_ctx* ctx, const epet_matrix* At, float alpha, const float* x,
                        float beta, float* y) {
  switch (At->value_type) {
    case EPET_VAL_F32: return bp_rows<ValF32>(ctx, At, alpha, x, beta, y);
    case EPET_VAL_BF16: return bp_rows<ValBF16>(ctx, At, alpha, x, beta, y);
    case EPET_VAL_F16: return bp_rows<ValF16>(ctx, At, alpha, x, beta, y);
  }
  EPET_FAIL(EPET_ERR_STATE, "unknown value type");
}

int matrix_spmv_basic(epet_ctx* ctx, const epet_matrix* A, int op, float alpha, const float* x,
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
