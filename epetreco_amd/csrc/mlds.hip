// mlds.hip -- multi-frame LDS-tiled products for batched MLEM.
//
// F frames share one matrix; images and data are frame-minor (X[v F + f]).
// The gathered vector is cut into tiles of TV = LDS_TILE / P columns, and a
// workgroup (tile t, frame block fb, row range g) of 1024 threads holds the P
// frames fb P .. fb P + P - 1 of its tile's columns in LDS (128 KB).  Its
// wavefronts walk R rows each over the row's entries inside the tile, taken
// from a tile-major copy of the matrix (16-bit tile-local columns), and
// accumulate P frames per entry from one or two 16-byte LDS reads.  Row r's
// partials of tile t go to partial[(t m + r) F + f]; reductions sum them in
// tile order (deterministic): the forward one fuses the MLEM ratio and the
// per-frame FP64 log-likelihood, the back-projection one writes the image.
// The matrix stream is read once per frame block (F / P times) instead of
// gathering F-float rows of X or E from the caches for every entry.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>

#include "epet_block.hpp"
#include "epet_internal.hpp"
#include "epet_rows.hpp"

namespace epet {

int build_tile_major(epet_ctx* ctx, const epet_matrix* A, int TV, ColMap mp, int* T_out,
                     int64_t** off_out, void** val_out, uint16_t** col_out);
ColMap voxel_map(const epet_ctx* ctx, const epet_matrix* A);

constexpr int MFL_BLOCK = 1024;
constexpr int MFL_WAVES = MFL_BLOCK / WAVE;

// TF: LDS floats per workgroup (32768 = 128 KB, one workgroup per CU; 16384 =
// 64 KB, two per CU with at most 64 VGPRs)
template <typename VT, int P, int R, int J, int BWD, int TF>
__device__ __forceinline__ void mfl_body(int m, int n, int F, int T, int FB, long long rows_per_wg,
                                         const int64_t* __restrict__ off,
                                         const uint16_t* __restrict__ col,
                                         const typename VT::T* __restrict__ val,
                                         const float* __restrict__ X,
                                         float* __restrict__ partial, ColMap mp) {
  extern __shared__ float xs[];
  constexpr int TV = TF / P;
  constexpr int L = WAVE / R;
  static_assert(P % 4 == 0, "frames per pass: multiple of 4");
  const int t = blockIdx.x % T;
  const long long rest = blockIdx.x / T;
  const int fb = (int)(rest % FB);
  const long long g = rest / FB;
  const long long tb = (long long)t * TV;
  const int cnt = (int)std::min<long long>(TV, (long long)n - tb);
  const int f0 = fb * P;
  // tile: xs[v P + f] = X[c F + f0 + f] for the tile's v-th column c (in the
// copy's column order), 0 for frames past F
  for (int i = threadIdx.x; i < cnt * P; i += MFL_BLOCK) {
    const int v = i / P, f = i % P;
    xs[i] = (f0 + f < F) ? X[colmap_inv(tb + v, mp) * F + f0 + f] : 0.0f;
  }
  __syncthreads();
  const int lane = lane_id(), gi = lane / L, l = lane % L;
  const int w = __builtin_amdgcn_readfirstlane(wave_id());
  const long long r0 = g * rows_per_wg;
  const long long r1 = std::min<long long>((long long)m, r0 + rows_per_wg);
  const unsigned last = (unsigned)(cnt - 1);
  const long long step = (long long)MFL_WAVES * R;
  const int64_t* offt = off + (long long)t * m;
  auto segment = [&](long long rbx, long long& s0, long long& s1) {
    s0 = 0;
    s1 = 0;
    const long long r = rbx + gi;
    if (r < r1) {
      s0 = offt[r];
      s1 = offt[r + 1];
    }
  };
  long long rb = r0 + (long long)w * R;
  long long n0, n1;
  segment(rb, n0, n1);
  for (; rb < r1; rb += step) {
    const long long r = rb + gi;
    const bool ok = r < r1;
    const long long s0 = n0, s1 = n1;
    segment(rb + step, n0, n1);
    float acc[P];
#pragma unroll
    for (int f = 0; f < P; f++) acc[f] = 0.0f;
    for (long long base = s0; base < s1; base += (long long)L * J) {
      uint16_t c[J];
      float v[J];
#pragma unroll
      for (int j = 0; j < J; j++) {
        const long long k = base + l + (long long)j * L;
        const bool in = k < s1;
        const long long kc = in ? k : s1 - 1;
        c[j] = col[kc];
        const float a = VT::get_cached(val, kc);
        v[j] = in ? a : 0.0f;
      }
#pragma unroll
      for (int j = 0; j < J; j++) {
        const unsigned li = std::min((unsigned)c[j], last);
        const float4* xp = reinterpret_cast<const float4*>(xs + li * P);
#pragma unroll
        for (int q = 0; q < P / 4; q++) {
          const float4 xv = xp[q];
          acc[4 * q + 0] += v[j] * xv.x;
          acc[4 * q + 1] += v[j] * xv.y;
          acc[4 * q + 2] += v[j] * xv.z;
          acc[4 * q + 3] += v[j] * xv.w;
        }
      }
    }
#pragma unroll
    for (int f = 0; f < P; f++)
#pragma unroll
      for (int o = L / 2; o > 0; o >>= 1) acc[f] += __shfl_xor(acc[f], o, L);
    // lane l of the row group stores the frames f with f mod L = l; empty
    // segments store nothing (the reductions skip them by their offsets)
    if (ok && s1 > s0) {
#pragma unroll
      for (int f = 0; f < P; f++)
        if (f % L == l && f0 + f < F) partial[((long long)t * m + r) * F + f0 + f] = acc[f];
    }
  }
}

template <typename VT, int P, int R, int J, int BWD, int TF>
__global__ __launch_bounds__(MFL_BLOCK) void k_mfl(int m, int n, int F, int T, int FB,
                                                   long long rows_per_wg,
                                                   const int64_t* __restrict__ off,
                                                   const uint16_t* __restrict__ col,
                                                   const typename VT::T* __restrict__ val,
                                                   const float* __restrict__ X,
                                                   float* __restrict__ partial, ColMap mp) {
  mfl_body<VT, P, R, J, BWD, TF>(m, n, F, T, FB, rows_per_wg, off, col, val, X, partial, mp);
}

// the 64 KB-tile variant: two workgroups per CU, so at most 64 VGPRs
template <typename VT, int P, int R, int J, int BWD, int TF>
__global__ __launch_bounds__(MFL_BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_mfl2(
    int m, int n, int F, int T, int FB, long long rows_per_wg, const int64_t* __restrict__ off,
    const uint16_t* __restrict__ col, const typename VT::T* __restrict__ val,
    const float* __restrict__ X, float* __restrict__ partial, ColMap mp) {
  mfl_body<VT, P, R, J, BWD, TF>(m, n, F, T, FB, rows_per_wg, off, col, val, X, partial, mp);
}

// Forward reduction: yhat = scale_r sum_t partial, E = Y / yhat (0 where
// yhat = 0; times the row scale when fold = 1), per-frame FP64
// log-likelihood summed per workgroup in LDS, one global atomic per frame.
__global__ __launch_bounds__(256) void k_mfl_ratio(int m, int F, int T,
                                                   const int64_t* __restrict__ off,
                                                   const float* __restrict__ partial,
                                                   const float* __restrict__ rowscale,
                                                   const float* __restrict__ Y,
                                                   float* __restrict__ E, int fold,
                                                   const int* __restrict__ itp,
                                                   double* __restrict__ loglik) {
  __shared__ double s_ll[64];
  if (threadIdx.x < 64) s_ll[threadIdx.x] = 0.0;
  __syncthreads();
  const long long tot = (long long)m * F;
  const long long stride_t = (long long)m * F;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (long long)gridDim.x * blockDim.x) {
    const long long r = i / F;
    const int f = (int)(i % F);
    float s = 0.0f;
    for (int t = 0; t < T; t++) {
      const int64_t* o = off + (long long)t * m + r;
      if (o[1] > o[0]) s += partial[(long long)t * stride_t + i];
    }
    const float sc = rowscale ? rowscale[r] : 1.0f;
    const float yh = s * sc;
    const float yi = Y[i];
    const float ev = yh > 0.0f ? yi / yh : 0.0f;
    E[i] = fold ? ev * sc : ev;
    if (yh > 0.0f) atomicAdd(&s_ll[f], (double)yi * log((double)yh) - (double)yh);
  }
  __syncthreads();
  if (loglik && threadIdx.x < F)
    unsafeAtomicAdd(&loglik[(long long)(*itp) * F + threadIdx.x], s_ll[threadIdx.x]);
}

// Back-projection reduction: C = sum_t partial over the non-empty segments
// (rows of the transposed copy)
__global__ __launch_bounds__(256) void k_mfl_sum(int m, int F, int T, const int64_t* __restrict__ off,
                                                 const float* __restrict__ partial,
                                                 float* __restrict__ C) {
  const long long tot = (long long)m * F;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= tot) return;
  const long long r = i / F;
  float s = 0.0f;
  for (int t = 0; t < T; t++) {
    const int64_t* o = off + (long long)t * m + r;
    if (o[1] > o[0]) s += partial[(long long)t * tot + i];
  }
  C[i] = s;
}

template <typename VT, int P, int R, int J, int BWD, int TF>
static void launch_mfl_t(epet_ctx* ctx, const epet_matrix* A, int F, const float* X) {
  auto kern = TF == LDS_TILE ? k_mfl<VT, P, R, J, BWD, TF> : k_mfl2<VT, P, R, J, BWD, TF>;
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              TF * (int)sizeof(float));
    attr = true;
  }
  const int FB = (F + P - 1) / P;
  const long long units = (long long)A->mfl_T * FB;
  // sixteen workgroups per CU in all, the fastest at 52^3 x 32 frames
  // (EPET_MFL_WG overrides it, for measurements)
  const char* ew = getenv("EPET_MFL_WG");
  const long long per_cu = ew ? std::max(1, atoi(ew)) : 16;
  const long long G = std::max<long long>(1, (per_cu * ctx->num_cus + units - 1) / units);
  const long long rows_per_wg = std::max<long long>(MFL_WAVES * R, (A->m + G - 1) / G);
  const long long ng = (A->m + rows_per_wg - 1) / rows_per_wg;
  kern<<<(unsigned)(ng * units), MFL_BLOCK, TF * sizeof(float), ctx->stream>>>(A->m, A->n, F, A->mfl_T, FB, rows_per_wg, A->mfl_off,
                                          A->mfl_col16, (const typename VT::T*)A->mfl_val, X,
                                          A->mfl_partial, A->mfl_map);
}

template <typename VT, int P, int R, int J, int BWD>
static void launch_mfl(epet_ctx* ctx, const epet_matrix* A, int F, const float* X) {
  if (A->mfl_tile == LDS_TILE / 2) launch_mfl_t<VT, P, R, J, BWD, LDS_TILE / 2>(ctx, A, F, X);
  else launch_mfl_t<VT, P, R, J, BWD, LDS_TILE>(ctx, A, F, X);
}

// rows per wavefront from the mean segment length (EPET_MFL_R = 4, 8 or 16
// forces it, for measurements)
template <typename VT, int BWD>
static void mfl_dispatch(epet_ctx* ctx, const epet_matrix* A, int F, const float* X) {
  const double seg = (double)A->nnz / std::max(1.0, (double)A->m * A->mfl_T);
  int R = seg >= 48.0 ? 4 : 16;
  const char* er = getenv("EPET_MFL_R");
  if (er) R = atoi(er);
  if (A->mfl_P == 8) {
    if (R == 4) launch_mfl<VT, 8, 4, 16, BWD>(ctx, A, F, X);
    else if (R == 8) launch_mfl<VT, 8, 8, 8, BWD>(ctx, A, F, X);
    else launch_mfl<VT, 8, 16, 8, BWD>(ctx, A, F, X);
  } else {
    if (R == 4) launch_mfl<VT, 4, 4, 16, BWD>(ctx, A, F, X);
    else if (R == 8) launch_mfl<VT, 4, 8, 8, BWD>(ctx, A, F, X);
    else launch_mfl<VT, 4, 16, 8, BWD>(ctx, A, F, X);
  }
}

template <int BWD>
static void mfl_product(epet_ctx* ctx, const epet_matrix* A, int F, const float* X) {
  switch (A->value_type) {
    case EPET_VAL_F32: mfl_dispatch<ValF32, BWD>(ctx, A, F, X); break;
    case EPET_VAL_BF16: mfl_dispatch<ValBF16, BWD>(ctx, A, F, X); break;
    default: mfl_dispatch<ValF16, BWD>(ctx, A, F, X); break;
  }
}

static bool mfl_wanted(const epet_matrix* A) {
  const char* e = getenv("EPET_MFL");
  if (e && e[0] == '0') return false;
  return A->m > 0 && A->cols_strict && (double)A->nnz / A->m >= 16.0;
}

// frames per pass for F frames (EPET_MFL_P = 4 or 8 forces it; EPET_MFL_P_BWD
// for the transposed copy alone, for measurements)
static int mfl_frames_per_pass(int F, bool bwd) {
  const char* eb = bwd ? getenv("EPET_MFL_P_BWD") : nullptr;
  if (eb) return atoi(eb) == 4 ? 4 : 8;
  const char* e = getenv("EPET_MFL_P");
  if (e) return atoi(e) == 4 ? 4 : 8;
  return F >= 8 ? 8 : 4;
}

static void mfl_release(epet_matrix* A) {
  if (A->mfl_off) (void)hipFree(A->mfl_off);
  if (A->mfl_val) (void)hipFree(A->mfl_val);
  if (A->mfl_col16) (void)hipFree(A->mfl_col16);
  if (A->mfl_partial) (void)hipFree(A->mfl_partial);
  A->mfl_off = nullptr;
  A->mfl_val = nullptr;
  A->mfl_col16 = nullptr;
  A->mfl_partial = nullptr;
  A->mfl_P = A->mfl_T = A->mfl_tile = 0;
  A->mfl_partial_len = 0;
}

// Tile-major copy for F frames and a partial buffer large enough (allocates
// and synchronises: call outside graph capture).  Without memory the matrix
// is left as it was and the gather kernels run.
static int mfl_prepare_one(epet_ctx* ctx, epet_matrix* A, int F, bool voxel_columns) {
  if (A->mfl_failed || !mfl_wanted(A)) return EPET_OK;
  const int P = mfl_frames_per_pass(F, !voxel_columns);
  // EPET_MFL_TILE_KB=64: 64 KB tiles, two workgroups per CU (measurements)
  const char* et = getenv("EPET_MFL_TILE_KB");
  const int tile = (et && atoi(et) == 64) ? LDS_TILE / 2 : LDS_TILE;
  if (A->mfl_off && (A->mfl_P != P || A->mfl_tile != tile)) mfl_release(A);
  if (!A->mfl_off) {
    int T = 0;
    const ColMap mp = voxel_columns ? voxel_map(ctx, A) : ColMap{};
    EPET_TRY(build_tile_major(ctx, A, tile / P, mp, &T, &A->mfl_off, &A->mfl_val,
                              &A->mfl_col16));
    A->mfl_map = mp;
    A->mfl_tile = tile;
    if (!A->mfl_off) {
      A->mfl_failed = true;
      return EPET_OK;
    }
    A->mfl_P = P;
    A->mfl_T = T;
  }
  const long long need = (long long)A->mfl_T * A->m * F;
  if (need > A->mfl_partial_len) {
    if (A->mfl_partial) (void)hipFree(A->mfl_partial);
    A->mfl_partial = nullptr;
    A->mfl_partial_len = 0;
    if (hipMalloc((void**)&A->mfl_partial, sizeof(float) * (size_t)std::max(1ll, need)) !=
        hipSuccess) {
      (void)hipGetLastError();
      mfl_release(A);
      A->mfl_failed = true;
      return EPET_OK;
    }
    A->mfl_partial_len = need;
  }
  return EPET_OK;
}

int mfl_prepare(epet_ctx* ctx, epet_matrix* A, int F) {
  if (!A->tr || F < 2) return EPET_OK;
  EPET_TRY(mfl_prepare_one(ctx, A, F, true));
  EPET_TRY(mfl_prepare_one(ctx, A->tr, F, false));
  return EPET_OK;
}

bool mfl_ready(const epet_matrix* A, int F) {
  const char* e = getenv("EPET_MFL");
  if (e && e[0] == '0') return false;
  auto ok = [F](const epet_matrix* M) {
    return M && M->mfl_off && M->mfl_partial_len >= (long long)M->mfl_T * M->m * F;
  };
  return F >= 2 && ok(A) && ok(A->tr);
}

// Forward + back projection of one multi-frame MLEM iteration (capturable once
// prepared): E from X, then C = A^T E.
int mfl_iteration(epet_ctx* ctx, const epet_matrix* A, const MfArgs& a) {
  const epet_matrix* At = A->tr;
  mfl_product<0>(ctx, A, a.F, a.x);
  k_mfl_ratio<<<grid_for((long long)A->m * a.F, 256, 16ll * ctx->num_cus), 256, 0, ctx->stream>>>(
      A->m, a.F, A->mfl_T, A->mfl_off, A->mfl_partial, A->rowscale, a.y, a.e, A->rowscale ? 1 : 0,
      a.itp, a.loglik);
  mfl_product<1>(ctx, At, a.F, a.e);
  const long long tot = (long long)At->m * a.F;
  k_mfl_sum<<<grid_for(tot, 256, 1ll << 31), 256, 0, ctx->stream>>>(At->m, a.F, At->mfl_T,
                                                                    At->mfl_off, At->mfl_partial,
                                                                    a.c);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

}  // namespace epet
