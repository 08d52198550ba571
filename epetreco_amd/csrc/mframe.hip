// mframe.hip -- multi-frame (batched) projection kernels and their autotuner.
//
// F frames share one system matrix; images and data are frame-minor
// (X[v*F + f]), so a matrix entry gathers one contiguous row of F floats.
// Both kernels walk matrix rows (forward: rows of A = channels; back
// projection: rows of the attached transpose = voxels) with
//   R rows per wavefront   (L = 64/R lanes per row),
//   P frames per lane tile (S = L/P entries of the row in flight per step),
// lane = (row g, slot s, frame f); a row's (column, value) pairs are loaded L at
// a time and broadcast inside the row's lane group with ds_bpermute.  (R, P)
// is picked per matrix and frame count by timing every valid pair
// (R in {1,2,4,8}, P in {4,8,16,32}, P <= L).
//
// Workgroup -> rows follows an XCD-aware mapping: workgroups are dispatched
// round-robin over the 8 XCDs, so workgroup b walks the contiguous eighth
// (b mod 8) of the rows and rows in flight on one XCD are neighbours whose
// image rows are shared in that XCD's L2 (locality only; correctness does not
// depend on the dispatch order).
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <vector>

#include "epet_block.hpp"
#include "epet_internal.hpp"
#include "epet_rows.hpp"

namespace epet {

constexpr int MF_NXCD = 8;
constexpr int MF_GATHERS = 16;   // gather registers per lane per round

// rows -> (first row group, stride, end) for this wavefront
__device__ __forceinline__ void mf_groups(long long ngroups, long long& g, long long& end,
                                          long long& stride) {
  const int xcd = blockIdx.x % MF_NXCD;
  const int kb = blockIdx.x / MF_NXCD;
  const int nb = (gridDim.x - xcd + MF_NXCD - 1) / MF_NXCD;
  const long long chunk = (ngroups + MF_NXCD - 1) / MF_NXCD;
  const long long beg = min(ngroups, xcd * chunk);
  end = min(ngroups, beg + chunk);
  g = beg + (long long)kb * (256 / WAVE) + wave_id();
  stride = (long long)nb * (256 / WAVE);
}

// acc[t] += sum over the row's entries [b, e) of v * G[col * F + t*P + f].
// A batch of L entries is loaded one (column, value) pair per lane, coalesced,
// one batch ahead of its use.  Each round then broadcasts QF steps of S
// entries inside the row's lane group and issues all their gathers
// (MF_GATHERS registers) before accumulating, so a wavefront keeps QF * nt gathers in
// flight.  Slots past the row end carry value 0 and a clamped column; frames
// past F read frame F-1 and are never stored.
template <typename VT, int L, int P, int MT>
__device__ __forceinline__ void mf_row(long long b, long long e, int l, int F, int nt,
                                       const int* __restrict__ col,
                                       const typename VT::T* __restrict__ val,
                                       const float* __restrict__ G, float acc[MT]) {
  constexpr int S = L / P;                       // entries per step
  constexpr int NS = P;                          // steps per batch
  constexpr int QF0 = MF_GATHERS / MT < 1 ? 1 : MF_GATHERS / MT;
  constexpr int QF = QF0 > NS ? NS : QF0;        // steps per round
  static_assert(NS % QF == 0, "round size");
  const int s = l / P, f = l % P;
  int fo[MT];
#pragma unroll
  for (int t = 0; t < MT; t++) fo[t] = min(t * P + f, F - 1);
  if (b >= e) return;
  // ordinary loads: with R > 1 a row group reads part of a cache line per load
  long long k = b + l;
  int ci = col[k < e ? k : e - 1];
  float cv = VT::get_cached(val, k < e ? k : e - 1);
  cv = k < e ? cv : 0.0f;
  for (long long base = b; base < e; base += L) {
    const long long kn = base + L + l;
    const long long knc = kn < e ? kn : e - 1;
    const int cin = col[knc];
    float cvn = VT::get_cached(val, knc);
    cvn = kn < e ? cvn : 0.0f;
#pragma unroll
    for (int q0 = 0; q0 < NS; q0 += QF) {
      float vq[QF];
      float g[QF][MT];
#pragma unroll
      for (int qq = 0; qq < QF; qq++) {
        const int src = (q0 + qq) * S + s;
        const int c = __shfl(ci, src, L);
        vq[qq] = __shfl(cv, src, L);
        const float* gp = G + (long long)c * F;
#pragma unroll
        for (int t = 0; t < MT; t++)
          if (t < nt) g[qq][t] = gp[fo[t]];
      }
#pragma unroll
      for (int qq = 0; qq < QF; qq++)
#pragma unroll
        for (int t = 0; t < MT; t++)
          if (t < nt) acc[t] += vq[qq] * g[qq][t];
    }
    ci = cin;
    cv = cvn;
  }
}

// lane-group sum over the S slots of each frame
template <int L, int P, int MT>
__device__ __forceinline__ void mf_reduce(float acc[MT]) {
#pragma unroll
  for (int t = 0; t < MT; t++)
#pragma unroll
    for (int off = L / 2; off >= P; off >>= 1) acc[t] += __shfl_xor(acc[t], off, L);
}

// Forward: E = Y / (scale * A X) (0 where the projection is 0), times the row
// scale when fold = 1; FP64 log-likelihood per frame, summed per workgroup in
// LDS and added to the global sum once per workgroup and frame.
template <typename VT, int R, int P>
__global__ __launch_bounds__(256) void k_mf_fwd(int m, int F, const int64_t* __restrict__ rowptr,
                                                const int* __restrict__ col,
                                                const typename VT::T* __restrict__ val,
                                                const float* __restrict__ rowscale,
                                                const float* __restrict__ X,
                                                const float* __restrict__ Y,
                                                float* __restrict__ E, int fold,
                                                const int* __restrict__ itp,
                                                double* __restrict__ loglik) {
  constexpr int L = WAVE / R;
  constexpr int MT = 64 / P;
  __shared__ double s_ll[64];
  if (threadIdx.x < 64) s_ll[threadIdx.x] = 0.0;
  __syncthreads();
  const int lane = lane_id();
  const int gi = lane / L, l = lane % L;
  const int s = l / P, f = l % P;
  const int nt = (F + P - 1) / P;
  const long long ngroups = ((long long)m + R - 1) / R;
  long long grp, end, stride;
  mf_groups(ngroups, grp, end, stride);
  for (; grp < end; grp += stride) {
    const long long r = grp * R + gi;
    const bool ok = r < m;
    const long long b = ok ? rowptr[r] : 0, e = ok ? rowptr[r + 1] : 0;
    float acc[MT];
#pragma unroll
    for (int t = 0; t < MT; t++) acc[t] = 0.0f;
    mf_row<VT, L, P, MT>(b, e, l, F, nt, col, val, X, acc);
    mf_reduce<L, P, MT>(acc);
    if (ok && s == 0) {
      const float sc = rowscale ? rowscale[r] : 1.0f;
#pragma unroll
      for (int t = 0; t < MT; t++) {
        const int fr = t * P + f;
        if (t < nt && fr < F) {
          const float yh = acc[t] * sc;
          const float yi = Y[r * F + fr];
          const float ev = yh > 0.0f ? yi / yh : 0.0f;
          E[r * F + fr] = fold ? ev * sc : ev;
          if (yh > 0.0f) atomicAdd(&s_ll[fr], (double)yi * log((double)yh) - (double)yh);
        }
      }
    }
  }
  __syncthreads();
  if (loglik && threadIdx.x < F) unsafeAtomicAdd(&loglik[(long long)(*itp) * F + threadIdx.x],
                                                 s_ll[threadIdx.x]);
}

// Back projection from the transpose: C[j][f] = sum_(i, v) v * E[i][f].
template <typename VT, int R, int P>
__global__ __launch_bounds__(256) void k_mf_bwd(int n, int F, const int64_t* __restrict__ tptr,
                                                const int* __restrict__ trow,
                                                const typename VT::T* __restrict__ tval,
                                                const float* __restrict__ E,
                                                float* __restrict__ C) {
  constexpr int L = WAVE / R;
  constexpr int MT = 64 / P;
  const int lane = lane_id();
  const int gi = lane / L, l = lane % L;
  const int s = l / P, f = l % P;
  const int nt = (F + P - 1) / P;
  const long long ngroups = ((long long)n + R - 1) / R;
  long long grp, end, stride;
  mf_groups(ngroups, grp, end, stride);
  for (; grp < end; grp += stride) {
    const long long j = grp * R + gi;
    const bool ok = j < n;
    const long long b = ok ? tptr[j] : 0, e = ok ? tptr[j + 1] : 0;
    float acc[MT];
#pragma unroll
    for (int t = 0; t < MT; t++) acc[t] = 0.0f;
    mf_row<VT, L, P, MT>(b, e, l, F, nt, trow, tval, E, acc);
    mf_reduce<L, P, MT>(acc);
    if (ok && s == 0) {
#pragma unroll
      for (int t = 0; t < MT; t++) {
        const int fr = t * P + f;
        if (t < nt && fr < F) C[j * F + fr] = acc[t];
      }
    }
  }
}

// ------------------------------------------------------------------ dispatch

template <typename VT, int R, int P>
static void mf_launch_fwd(epet_ctx* ctx, const epet_matrix* A, const MfArgs& a) {
  using T = typename VT::T;
  const long long ng = ((long long)A->m + R - 1) / R;
  const unsigned g = (unsigned)std::max<long long>(MF_NXCD, std::min<long long>((ng + 3) / 4, 16ll * ctx->num_cus));
  k_mf_fwd<VT, R, P><<<g, 256, 0, ctx->stream>>>(A->m, a.F, A->rowptr, A->col, (const T*)A->val,
                                                 A->rowscale, a.x, a.y, a.e,
                                                 A->rowscale ? 1 : 0, a.itp, a.loglik);
}

template <typename VT, int R, int P>
static void mf_launch_bwd(epet_ctx* ctx, const epet_matrix* At, const MfArgs& a) {
  using T = typename VT::T;
  const long long ng = ((long long)At->m + R - 1) / R;
  const unsigned g = (unsigned)std::max<long long>(MF_NXCD, std::min<long long>((ng + 3) / 4, 16ll * ctx->num_cus));
  k_mf_bwd<VT, R, P><<<g, 256, 0, ctx->stream>>>(At->m, a.F, At->rowptr, At->col,
                                                 (const T*)At->val, a.e, a.c);
}

// every (R, P) pair of the tuning space
static const int MF_PAIRS[][2] = {{1, 4}, {1, 8}, {1, 16}, {1, 32}, {2, 4}, {2, 8}, {2, 16},
                                  {2, 32}, {4, 4}, {4, 8}, {4, 16}, {8, 4}, {8, 8}};
constexpr int MF_NPAIRS = sizeof(MF_PAIRS) / sizeof(MF_PAIRS[0]);

template <typename VT>
static bool mf_dispatch(epet_ctx* ctx, const epet_matrix* A, bool fwd, int R, int P,
                        const MfArgs& a) {
  const epet_matrix* M = fwd ? A : A->tr;
#define EPET_MF(RR, PP)                                                  \
  if (R == RR && P == PP) {                                              \
    if (fwd) mf_launch_fwd<VT, RR, PP>(ctx, M, a);                       \
    else mf_launch_bwd<VT, RR, PP>(ctx, M, a);                           \
    return true;                                                         \
  }
  EPET_MF(1, 4) EPET_MF(1, 8) EPET_MF(1, 16) EPET_MF(1, 32) EPET_MF(1, 64)
  EPET_MF(2, 4) EPET_MF(2, 8) EPET_MF(2, 16) EPET_MF(2, 32)
  EPET_MF(4, 4) EPET_MF(4, 8) EPET_MF(4, 16)
  EPET_MF(8, 4) EPET_MF(8, 8)
#undef EPET_MF
  return false;
}

static bool mf_call(epet_ctx* ctx, const epet_matrix* A, bool fwd, int R, int P, const MfArgs& a) {
  switch (A->value_type) {
    case EPET_VAL_F32: return mf_dispatch<ValF32>(ctx, A, fwd, R, P, a);
    case EPET_VAL_BF16: return mf_dispatch<ValBF16>(ctx, A, fwd, R, P, a);
    default: return mf_dispatch<ValF16>(ctx, A, fwd, R, P, a);
  }
}

static void mf_default(int F, int& R, int& P) {
  R = 1;
  P = F <= 4 ? 4 : F <= 8 ? 8 : F <= 16 ? 16 : F <= 32 ? 32 : 64;
}

// Times every valid pair (one warm-up and `reps` timed launches each, HIP
// events on the context's stream) and keeps the fastest forward and back
// projection configuration for frame count F.
int mf_tune(epet_ctx* ctx, epet_matrix* A, const MfArgs& a, int reps) {
  EPET_CHECK(A->tr, EPET_ERR_STATE, "multi-frame tuning needs the transposed copy");
  hipEvent_t e0, e1;
  EPET_HIP(hipEventCreate(&e0));
  EPET_HIP(hipEventCreate(&e1));
  MfArgs b = a;
  b.loglik = nullptr;
  for (int which = 0; which < 2; which++) {
    const bool fwd = which == 0;
    float best = 1e30f;
    int bR, bP;
    mf_default(a.F, bR, bP);
    for (int p = 0; p < MF_NPAIRS; p++) {
      const int R = MF_PAIRS[p][0], P = MF_PAIRS[p][1];
      if (P > 4 && P / 2 >= a.F) continue;   // tile wider than the frames
      mf_call(ctx, A, fwd, R, P, b);
      EPET_HIP(hipEventRecord(e0, ctx->stream));
      for (int k = 0; k < reps; k++) mf_call(ctx, A, fwd, R, P, b);
      EPET_HIP(hipEventRecord(e1, ctx->stream));
      EPET_HIP(hipEventSynchronize(e1));
      EPET_HIP(hipGetLastError());
      float ms = 0.0f;
      EPET_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) { best = ms; bR = R; bP = P; }
    }
    if (fwd) { A->mf_fwd_R = bR; A->mf_fwd_P = bP; A->mf_fwd_ms = best / reps; }
    else { A->mf_bwd_R = bR; A->mf_bwd_P = bP; A->mf_bwd_ms = best / reps; }
  }
  A->mf_F = a.F;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return EPET_OK;
}

// Forward + back projection of one multi-frame MLEM iteration (capturable).
int mf_iteration(epet_ctx* ctx, const epet_matrix* A, const MfArgs& a) {
  int Rf = A->mf_fwd_R, Pf = A->mf_fwd_P, Rb = A->mf_bwd_R, Pb = A->mf_bwd_P;
  if (A->mf_F != a.F) { mf_default(a.F, Rf, Pf); Rb = Rf; Pb = Pf; }
  EPET_CHECK(mf_call(ctx, A, true, Rf, Pf, a), EPET_ERR_UNSUPPORTED, "no multi-frame kernel");
  EPET_CHECK(mf_call(ctx, A, false, Rb, Pb, a), EPET_ERR_UNSUPPORTED, "no multi-frame kernel");
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

}  // namespace epet
