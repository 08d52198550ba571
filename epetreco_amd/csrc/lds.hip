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
// scale, alpha/beta or the MLEM ratio.
//
// The kernels read a tile-major copy of the matrix built once per matrix: the
// entries of (tile t, row r) lie contiguously at lds_off[t m + r], rows of one
// tile back to back, columns stored tile-local in 16 bits.  A cache line that
// straddles two segments then belongs to neighbouring rows of the same
// workgroup and is read from HBM once, and an FP32 entry costs 6 bytes.  HBM
// sees the entry stream once, x once per workgroup, and 12 bytes of offset and
// partial per (row, tile).
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

// A chunk of LDS_CHUNK entries of the tile-major copy: eight 16-bit columns
// (one 16-byte load) and eight values (one or two 16-byte loads).
template <typename VT> struct LdsChunk;
template <> struct LdsChunk<ValF32> {
  static __device__ __forceinline__ void load(const float* __restrict__ val, long long k,
                                              float v[8]) {
    const float4 a = *reinterpret_cast<const float4*>(val + k);
    const float4 b = *reinterpret_cast<const float4*>(val + k + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  }
};
template <> struct LdsChunk<ValBF16> {
  static __device__ __forceinline__ void load(const uint16_t* __restrict__ val, long long k,
                                              float v[8]) {
    const uint4 a = *reinterpret_cast<const uint4*>(val + k);
    const unsigned w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      v[2 * i] = __uint_as_float(w[i] << 16);
      v[2 * i + 1] = __uint_as_float(w[i] & 0xFFFF0000u);
    }
  }
};
template <> struct LdsChunk<ValF16> {
  static __device__ __forceinline__ void load(const uint16_t* __restrict__ val, long long k,
                                              float v[8]) {
    const uint4 a = *reinterpret_cast<const uint4*>(val + k);
    const unsigned w[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      v[2 * i] = __half2float(__ushort_as_half((uint16_t)(w[i] & 0xFFFFu)));
      v[2 * i + 1] = __half2float(__ushort_as_half((uint16_t)(w[i] >> 16)));
    }
  }
};

// R rows per wavefront (L = 64/R lanes per row), J chunks of LDS_CHUNK entries
// per lane per pass; BWD = 1 marks the back projection over a transposed copy
// (a separate instantiation, so that profiles tell the two products apart).
template <typename VT, int R, int J, int BWD>
__global__ __launch_bounds__(LDS_BLOCK) void k_rows_lds(int m, int n, int T, long long rows_per_wg,
                                                        const int64_t* __restrict__ off,
                                                        const uint16_t* __restrict__ col,
                                                        const typename VT::T* __restrict__ val,
                                                        const float* __restrict__ x,
                                                        float* __restrict__ partial, ColMap mp) {
  extern __shared__ float xs[];
  constexpr int L = WAVE / R;
  const int t = blockIdx.x % T;
  const long long g = blockIdx.x / T;
  const long long tb = (long long)t * LDS_TILE;
  const int cnt = (int)std::min<long long>(LDS_TILE, (long long)n - tb);
  if (mp.nx == 0) {
    // tile copy: 16-byte loads when the tile start is 16-byte aligned
    const int c4 = ((reinterpret_cast<uintptr_t>(x + tb) & 15) == 0) ? (cnt >> 2) : 0;
    for (int i = threadIdx.x; i < c4; i += LDS_BLOCK)
      reinterpret_cast<float4*>(xs)[i] = reinterpret_cast<const float4*>(x + tb)[i];
    for (int i = (c4 << 2) + threadIdx.x; i < cnt; i += LDS_BLOCK) xs[i] = x[tb + i];
  } else {
    // renumbered columns: tile entry i is column colmap_inv(tb + i) of x
    for (int i = threadIdx.x; i < cnt; i += LDS_BLOCK) xs[i] = x[colmap_inv(tb + i, mp)];
  }
  __syncthreads();
  const int lane = lane_id(), gi = lane / L, l = lane % L;
  const int w = __builtin_amdgcn_readfirstlane(wave_id());
  const long long r0 = g * rows_per_wg;
  const long long r1 = std::min<long long>((long long)m, r0 + rows_per_wg);
  const unsigned last = (unsigned)(cnt - 1);
  const long long step = (long long)LDS_WAVES * R;
  const int64_t* offt = off + (long long)t * m;
  // segment [s0, s1) of row rbx + gi in tile t (whole chunks); empty past the range
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
    segment(rb + step, n0, n1);   // the next segment's bounds, requested before this one's entries
    float acc = 0.0f;
    for (long long base = s0 + (long long)l * LDS_CHUNK; base < s1;
         base += (long long)L * LDS_CHUNK * J) {
      // J chunks: columns and values requested together (ordinary cached
      // loads), chunks past the segment skipped
      uint4 c[J];
      float v[J][LDS_CHUNK];
      bool in[J];
#pragma unroll
      for (int j = 0; j < J; j++) {
        const long long k = base + (long long)j * L * LDS_CHUNK;
        in[j] = k < s1;
        const long long kc = in[j] ? k : base;
        c[j] = *reinterpret_cast<const uint4*>(col + kc);
        LdsChunk<VT>::load(val, kc, v[j]);
      }
#pragma unroll
      for (int j = 0; j < J; j++) {
        const unsigned cw[4] = {c[j].x, c[j].y, c[j].z, c[j].w};
        float a = 0.0f;
#pragma unroll
        for (int i = 0; i < 4; i++) {
          // the clamp keeps LDS reads in range whatever the data
          const unsigned lo = std::min(cw[i] & 0xFFFFu, last), hi = std::min(cw[i] >> 16, last);
          a += v[j][2 * i] * xs[lo];
          a += v[j][2 * i + 1] * xs[hi];
        }
        acc += in[j] ? a : 0.0f;
      }
    }
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, L);
    if (ok && l == 0) partial[(long long)t * m + r] = acc;
  }
}

// ------------------------------------------------------------------ tile-major copy
// Stable counting sort of every row's entries by the voxel's y index (one
// wavefront per row, per-wave bucket counters in LDS): rows ascending in
// (ix, iy, iz) become ascending in (iy, ix, iz), i.e. in the y-major column
// colmap_fwd(c).  Writes the renumbered columns and the values in that order.
template <typename T>
__global__ __launch_bounds__(256) void k_ysort_rows(int m, const int64_t* __restrict__ rowptr,
                                                    const int* __restrict__ col,
                                                    const T* __restrict__ val, ColMap mp,
                                                    int* __restrict__ colp, T* __restrict__ valp) {
  extern __shared__ int bucket[];                 // (256 / WAVE) x ny
  const int lane = lane_id();
  const int w = wave_id();
  int* cnt = bucket + w * mp.ny;
  const long long nw = (long long)gridDim.x * (256 / WAVE);
  for (long long r = (long long)blockIdx.x * (256 / WAVE) + w; r < m; r += nw) {
    const long long b = rowptr[r], e = rowptr[r + 1];
    for (int i = lane; i < mp.ny; i += WAVE) cnt[i] = 0;
    __builtin_amdgcn_wave_barrier();
    for (long long k = b + lane; k < e; k += WAVE) atomicAdd(&cnt[(col[k] / mp.nz) % mp.ny], 1);
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
      int run = 0;
      for (int i = 0; i < mp.ny; i++) {
        const int c = cnt[i];
        cnt[i] = run;
        run += c;
      }
    }
    __builtin_amdgcn_wave_barrier();
    for (long long base = b; base < e; base += WAVE) {
      const long long k = base + lane;
      bool act = k < e;
      const int c = act ? col[k] : 0;
      const int iy = act ? (c / mp.nz) % mp.ny : -1;
      // lanes grouped by y, each group placed after the earlier entries of its
      // bucket in lane (= entry) order
      unsigned long long left = __ballot(act);
      while (left) {
        const int lead = __ffsll((long long)left) - 1;
        const int key = __shfl(iy, lead);
        const unsigned long long mask = __ballot(act && iy == key);
        const int start = cnt[key];
        if (act && iy == key) {
          const int rank = __popcll(mask & ((1ull << lane) - 1ull));
          colp[b + start + rank] = (int)colmap_fwd(c, mp);
          valp[b + start + rank] = val[k];
          act = false;
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == lead) cnt[key] = start + __popcll(mask);
        __builtin_amdgcn_wave_barrier();
        left &= ~mask;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// padded segment lengths: every (tile, row) segment is a whole number of
// LDS_CHUNK-entry chunks, so a lane reads a chunk's columns and values with
// 16-byte loads
__global__ __launch_bounds__(256) void k_lds_counts(int m, int T, const uint32_t* __restrict__ toff,
                                                    int* __restrict__ cnt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;   // i = t m + r
  if (i >= (long long)m * T) return;
  const int t = (int)(i / m);
  const long long r = i % m;
  const uint32_t* o = toff + r * (T + 1) + t;
  cnt[i] = ((int)(o[1] - o[0]) + LDS_CHUNK - 1) & ~(LDS_CHUNK - 1);
}

__global__ void k_lds_set_last(int64_t* __restrict__ off, long long i, long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) off[i] = v;
}

// one wavefront per (tile, row) segment: copy values and tile-local columns,
// then zero entries (column 0, value 0) up to the padded length
template <typename T>
__global__ __launch_bounds__(256) void k_lds_fill(int m, int T_, int mask,
                                                  const int64_t* __restrict__ rowptr,
                                                  const uint32_t* __restrict__ toff,
                                                  const int* __restrict__ col,
                                                  const T* __restrict__ val,
                                                  const int64_t* __restrict__ off,
                                                  uint16_t* __restrict__ col16,
                                                  T* __restrict__ val2) {
  const int lane = lane_id();
  const long long nw = (long long)gridDim.x * (256 / WAVE);
  for (long long i = (long long)blockIdx.x * (256 / WAVE) + wave_id(); i < (long long)m * T_;
       i += nw) {
    const int t = (int)(i / m);
    const long long r = i % m;
    const uint32_t* o = toff + r * (T_ + 1) + t;
    const long long src = rowptr[r] + o[0];
    const long long len = o[1] - o[0];
    const long long dst = off[i], plen = off[i + 1] - dst;
    for (long long k = lane; k < plen; k += WAVE) {
      const bool in = k < len;
      col16[dst + k] = in ? (uint16_t)(col[src + k] & mask) : (uint16_t)0;
      val2[dst + k] = in ? val[src + k] : (T)0;
    }
  }
}

// ------------------------------------------------------------------ reductions
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

// ------------------------------------------------------------------ dispatch
template <typename VT, int R, int J, int BWD>
static void launch_rows_lds(epet_ctx* ctx, const epet_matrix* A, const float* x) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)k_rows_lds<VT, R, J, BWD>,
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              LDS_TILE * (int)sizeof(float));
    attr = true;
  }
  // each workgroup owns its CU's LDS; several workgroups per CU in all, so
  // that tiles and row ranges of unequal weight even out: 12 for the forward
  // and 24 for the back projection at 52^3 (4 to 32 per CU were within a few
  // percent of each other)
  const long long per_cu = BWD ? 24 : 12;
  const long long G = std::max<long long>(1, (per_cu * ctx->num_cus + A->lds_T - 1) / A->lds_T);
  const long long rows_per_wg = std::max<long long>(LDS_WAVES * R, (A->m + G - 1) / G);
  const long long ng = (A->m + rows_per_wg - 1) / rows_per_wg;
  k_rows_lds<VT, R, J, BWD><<<(unsigned)(ng * A->lds_T), LDS_BLOCK, LDS_TILE * sizeof(float),
                             ctx->stream>>>(A->m, A->n, A->lds_T, rows_per_wg, A->lds_off,
                                            A->lds_col16, (const typename VT::T*)A->lds_val, x,
                                            A->lds_partial, A->lds_map);
}

// Rows per wavefront from the mean segment length; one chunk per lane per pass
// (at 52^3 two chunks per pass took 8.1 ms per MLEM iteration, one 6.9-7.2 ms).
template <typename VT, int BWD>
static void rows_lds_dispatch(epet_ctx* ctx, const epet_matrix* A, const float* x) {
  const double seg = (double)A->nnz / std::max(1.0, (double)A->m * A->lds_T);
  if (seg >= 384.0) launch_rows_lds<VT, 2, 1, BWD>(ctx, A, x);
  else if (seg >= 48.0) launch_rows_lds<VT, 4, 1, BWD>(ctx, A, x);
  else launch_rows_lds<VT, 16, 1, BWD>(ctx, A, x);
}

template <int BWD>
static void rows_lds(epet_ctx* ctx, const epet_matrix* A, const float* x) {
  switch (A->value_type) {
    case EPET_VAL_F32: rows_lds_dispatch<ValF32, BWD>(ctx, A, x); break;
    case EPET_VAL_BF16: rows_lds_dispatch<ValBF16, BWD>(ctx, A, x); break;
    default: rows_lds_dispatch<ValF16, BWD>(ctx, A, x); break;
  }
}

// Whether the LDS-tiled products apply (the context may turn them off): long
// rows, ascending columns.
static bool lds_wanted(const epet_ctx* ctx, const epet_matrix* A) {
  return ctx->opt.lds && A->m > 0 && A->cols_strict && (double)A->nnz / A->m >= 64.0;
}

static void lds_release(epet_matrix* A) {
  if (A->lds_off) (void)hipFree(A->lds_off);
  if (A->lds_val) (void)hipFree(A->lds_val);
  if (A->lds_col16) (void)hipFree(A->lds_col16);
  if (A->lds_partial) (void)hipFree(A->lds_partial);
  A->lds_off = nullptr;
  A->lds_val = nullptr;
  A->lds_col16 = nullptr;
  A->lds_partial = nullptr;
  A->lds_T = 0;
}

// Tile-major copy of A for column tiles of TV columns (a power of two up to
// 65536): *T_out tiles, offsets (T m + 1), values and tile-local 16-bit
// columns.  Allocates and synchronises.  On an allocation failure returns
// EPET_OK with *off == nullptr.
int build_tile_major(epet_ctx* ctx, const epet_matrix* A, int TV, ColMap mp, int* T_out,
                     int64_t** off_out, void** val_out, uint16_t** col_out) {
  *off_out = nullptr;
  *val_out = nullptr;
  *col_out = nullptr;
  hipStream_t st = ctx->stream;
  const int* scol = A->col;
  const void* sval = A->val;
  int* colp = nullptr;
  void* valp = nullptr;
  if (mp.nx > 0) {
    // renumbered columns: rows re-sorted into y-major column order first
    const size_t vb0 = A->value_type == EPET_VAL_F32 ? 4 : 2;
    if (hipMalloc((void**)&colp, sizeof(int) * (size_t)std::max(1ll, A->nnz)) != hipSuccess ||
        hipMalloc(&valp, vb0 * (size_t)std::max(1ll, A->nnz)) != hipSuccess) {
      (void)hipGetLastError();
      if (colp) (void)hipFree(colp);
      return EPET_OK;
    }
    const unsigned g = (unsigned)std::max<long long>(
        1, std::min<long long>(((long long)A->m + 3) / 4, 16ll * ctx->num_cus));
    const size_t lds = sizeof(int) * (256 / WAVE) * (size_t)mp.ny;
    if (vb0 == 4)
      k_ysort_rows<float><<<g, 256, lds, st>>>(A->m, A->rowptr, A->col, (const float*)A->val, mp,
                                               colp, (float*)valp);
    else
      k_ysort_rows<uint16_t><<<g, 256, lds, st>>>(A->m, A->rowptr, A->col,
                                                  (const uint16_t*)A->val, mp, colp,
                                                  (uint16_t*)valp);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
      (void)hipFree(colp);
      (void)hipFree(valp);
      EPET_FAIL(EPET_ERR_HIP, std::string("k_ysort_rows: ") + hipGetErrorString(le));
    }
    scol = colp;
    sval = valp;
  }
  const int T = (int)(((long long)A->n + TV - 1) / TV);
  const long long nseg = (long long)A->m * T;
  const size_t vb = A->value_type == EPET_VAL_F32 ? 4 : 2;
  uint32_t* toff = nullptr;
  int* cnt = nullptr;
  int64_t* off = nullptr;
  void* val = nullptr;
  uint16_t* col16 = nullptr;
  bool ok = T >= 1 &&
            hipMalloc((void**)&toff, sizeof(uint32_t) * (size_t)A->m * (T + 1)) == hipSuccess &&
            hipMalloc((void**)&cnt, sizeof(int) * (size_t)std::max(1ll, nseg)) == hipSuccess &&
            hipMalloc((void**)&off, sizeof(int64_t) * (size_t)(nseg + 1)) == hipSuccess;
  int rc = EPET_OK;
  if (ok) {
    launch_row_plan(st, A->m, T, TV, A->rowptr, scol, toff);
    k_lds_counts<<<grid_for(nseg, 256, 1ll << 31), 256, 0, st>>>(A->m, T, toff, cnt);
    long long tot = 0;
    rc = exclusive_scan_i32_to_i64(ctx, cnt, off, nseg, &tot);
    if (rc == EPET_OK && (tot < A->nnz || tot > A->nnz + (long long)(LDS_CHUNK - 1) * nseg))
      rc = fail(EPET_ERR_STATE, __FILE__, __LINE__, "tile-major copy: entry count mismatch");
    ok = rc == EPET_OK &&
         hipMalloc(&val, vb * (size_t)std::max(1ll, tot)) == hipSuccess &&
         hipMalloc((void**)&col16, sizeof(uint16_t) * (size_t)std::max(1ll, tot)) == hipSuccess;
    if (rc == EPET_OK && !ok) (void)hipGetLastError();
    if (ok) {
      k_lds_set_last<<<1, 64, 0, st>>>(off, nseg, tot);
      const unsigned g = (unsigned)std::max<long long>(
          1, std::min<long long>((nseg + 3) / 4, 16ll * ctx->num_cus));
      if (vb == 4)
        k_lds_fill<float><<<g, 256, 0, st>>>(A->m, T, TV - 1, A->rowptr, toff, scol,
                                             (const float*)sval, off, col16, (float*)val);
      else
        k_lds_fill<uint16_t><<<g, 256, 0, st>>>(A->m, T, TV - 1, A->rowptr, toff, scol,
                                                (const uint16_t*)sval, off, col16,
                                                (uint16_t*)val);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "tile-major copy");
    }
  } else {
    (void)hipGetLastError();
  }
  if (toff) (void)hipFree(toff);
  if (cnt) (void)hipFree(cnt);
  if (colp) (void)hipFree(colp);
  if (valp) (void)hipFree(valp);
  if (!ok || rc != EPET_OK) {
    if (off) (void)hipFree(off);
    if (val) (void)hipFree(val);
    if (col16) (void)hipFree(col16);
    return rc;
  }
  *T_out = T;
  *off_out = off;
  *val_out = val;
  *col_out = col16;
  return EPET_OK;
}

// y-major column order for the copy of a grid matrix when the context asks
// for it (EPET_LDS_MAP=1).  Off by default: rows then meet fewer tiles, but
// their segments become very uneven (most empty, a few long), which leaves the
// row groups of a wavefront idle; at 52^3 the forward product took 5.6 ms
// instead of 4.5 ms.
static ColMap voxel_map(const epet_ctx* ctx, const epet_matrix* A) {
  ColMap mp;
  if (!ctx->opt.lds_map || !ctx->have_geo) return mp;
  const int* gn = ctx->geo.gn;
  if ((long long)gn[0] * gn[1] * gn[2] != A->n || gn[1] > 4096) return mp;
  mp.nx = gn[0];
  mp.ny = gn[1];
  mp.nz = gn[2];
  return mp;
}

// Builds the tile-major copy and the partial buffer once (allocates and
// synchronises: call outside graph capture).  voxel_columns: the columns are
// image voxels (a forward projector), so they may be renumbered.  Leaves the
// matrix without them when an allocation fails; the other products then run.
int lds_prepare_map(epet_ctx* ctx, epet_matrix* A, bool voxel_columns) {
  if (A->lds_off || A->lds_failed || !lds_wanted(ctx, A)) return EPET_OK;
  int T = 0;
  const ColMap mp = voxel_columns ? voxel_map(ctx, A) : ColMap{};
  EPET_TRY(build_tile_major(ctx, A, LDS_TILE, mp, &T, &A->lds_off, &A->lds_val, &A->lds_col16));
  A->lds_map = mp;
  if (A->lds_off && hipMalloc((void**)&A->lds_partial, sizeof(float) * (size_t)A->m * T) ==
                        hipSuccess) {
    A->lds_T = T;
    return EPET_OK;
  }
  (void)hipGetLastError();
  lds_release(A);
  A->lds_failed = true;
  return EPET_OK;
}

int lds_prepare(epet_ctx* ctx, epet_matrix* A) { return lds_prepare_map(ctx, A, false); }

bool lds_ready(const epet_ctx* ctx, const epet_matrix* A) {
  return A->lds_off != nullptr && lds_wanted(ctx, A);
}

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
