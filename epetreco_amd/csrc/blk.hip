// blk.hip -- batched (multi-frame) projections for MLEM with the operator
// streamed once per iteration and direction.
//
// F frames share one matrix M (m rows, n columns); the gathered vector G and
// the output are frame-minor (G[c F + f]).  The matrix is kept once more as a
// blocked copy:
//   * rows in a locality order (rowperm: position -> row) and columns in a
//     locality order (colinv: position -> column), both cut into pieces: row
//     blocks of RB positions (chosen at launch, a multiple of BLK_G) and column
//     tiles of C positions (C Fp floats = 128 KB of LDS, Fp = padded frames);
//   * entries tile-major: for each tile, the segments of all row positions back
//     to back, each padded with zero entries to a multiple of 4 (16-byte loads);
//     an entry is (tile-local column, stored value), packed in 32 bits for
//     16-bit values;
//   * cnt[h m_pad + p]: padded length of row position p's segment in tile h;
//     seg_base[h (nG+1) + j]: first entry of granule j (BLK_G row positions) in
//     tile h, so that a row block's entries in a tile are one contiguous range.
//
// A workgroup owns a row block and keeps its rows' F-frame sums in registers
// while it walks the column tiles: for every tile the block touches it stages
// the tile's C columns x Fp frames of G in LDS, and each lane accumulates P
// frames of K rows (L = Fp / P lanes per row).  The matrix streams once, no
// partial sums per (row, tile) reach memory, and the row's complete sums feed
// the epilogue: the MLEM ratio and the FP64 log-likelihood for the forward
// projection, the image for the back projection.  When a direction has few
// row blocks (the back projection's voxel rows), each block's tiles are split
// by entry count over `splits` workgroups that write partial sums, which a
// reduction adds in split order.  The grid is persistent (one workgroup per CU
// looping over the work items), and the (K, P) shape is autotuned per matrix.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "epet_block.hpp"
#include "epet_internal.hpp"
#include "epet_rows.hpp"

namespace epet {

constexpr int BLK_THREADS = 1024;
constexpr int BLK_WAVES = BLK_THREADS / WAVE;
constexpr int BLK_TILE_FLOATS = 32768;      // 128 KB of the gathered vector per tile
constexpr int BLK_G = 64;                   // row granule of the segment bases
constexpr int BLK_MAX_RB = 4 * BLK_THREADS; // rows per block (one scan pass)

// Blocked copy of one matrix for one padded frame count.
struct BlkCopy {
  int Fp = 0, lgFp = 0;        // padded frames (power of two, >= 8)
  int C = 0, lgC = 0, T = 0;   // tile columns, tiles
  int m = 0, m_pad = 0, nG = 0;
  int vt = EPET_VAL_F32;
  int* rowperm = nullptr;      // m_pad: position -> row, -1 past m
  int* colinv = nullptr;       // T C: position -> column, -1 past n
  int64_t* seg_base = nullptr; // T (nG + 1)
  uint16_t* cnt = nullptr;     // T m_pad
  void* ent = nullptr;         // u32 packed (16-bit values) or u16 columns (FP32)
  float* eval = nullptr;       // FP32 values (FP32 storage)
  long long nent = 0;
  float* partial = nullptr;    // splits x m_pad x Fp
  long long partial_len = 0;
  // launch shape (autotuned): frames per lane, row groups per lane, splits
  int P = 4, K = 4, splits = 1;
  int tuned_F = 0;
  float ms = 0.0f;
  BlkCopy* next = nullptr;     // copies for other padded frame counts
  ~BlkCopy() {
    if (rowperm) (void)hipFree(rowperm);
    if (colinv) (void)hipFree(colinv);
    if (seg_base) (void)hipFree(seg_base);
    if (cnt) (void)hipFree(cnt);
    if (ent) (void)hipFree(ent);
    if (eval) (void)hipFree(eval);
    if (partial) (void)hipFree(partial);
    delete next;
  }
};

void blk_release(epet_matrix* A) {
  delete A->blk;
  A->blk = nullptr;
}

struct BlkArgs {
  int F, Fp, lgFp, C, T, m_pad, nG;
  int rb, nblocks, splits;
  int L, S;                    // lanes per row, row slots per wavefront
  int mode;                    // 0: forward MLEM ratio, 1: image, 2: partial sums
  int F4;                      // G rows are float4-aligned (F % 4 == 0)
  const int* rowperm;
  const int* colinv;
  const int64_t* seg_base;
  const uint16_t* cnt;
  const void* ent;
  const float* eval;
  const float* G;              // gathered vector, n x F (frame-minor)
  float* out;                  // forward: E (m x F); back projection: C (n x F)
  float* partial;
  const float* rowscale;       // forward: per-row scale of the stored values
  int fold;                    // forward: E carries the row scale
  const float* Y;              // forward: data (m x F)
  const int* itp;
  double* loglik;
};

// ------------------------------------------------------------------ entries
// Four consecutive entries (16-byte aligned): tile-local columns (masked to the
// tile, which also bounds every LDS index) and values.
template <typename VT> struct BlkEnt;
template <> struct BlkEnt<ValBF16> {
  static __device__ __forceinline__ void load(const BlkArgs& a, long long e, unsigned cm,
                                              unsigned c[4], float v[4]) {
    const uint4 pk = *reinterpret_cast<const uint4*>(static_cast<const uint32_t*>(a.ent) + e);
    c[0] = pk.x & cm; c[1] = pk.y & cm; c[2] = pk.z & cm; c[3] = pk.w & cm;
    v[0] = __uint_as_float(pk.x & 0xFFFF0000u);
    v[1] = __uint_as_float(pk.y & 0xFFFF0000u);
    v[2] = __uint_as_float(pk.z & 0xFFFF0000u);
    v[3] = __uint_as_float(pk.w & 0xFFFF0000u);
  }
};
template <> struct BlkEnt<ValF16> {
  static __device__ __forceinline__ void load(const BlkArgs& a, long long e, unsigned cm,
                                              unsigned c[4], float v[4]) {
    const uint4 pk = *reinterpret_cast<const uint4*>(static_cast<const uint32_t*>(a.ent) + e);
    c[0] = pk.x & cm; c[1] = pk.y & cm; c[2] = pk.z & cm; c[3] = pk.w & cm;
    v[0] = __half2float(__ushort_as_half((uint16_t)(pk.x >> 16)));
    v[1] = __half2float(__ushort_as_half((uint16_t)(pk.y >> 16)));
    v[2] = __half2float(__ushort_as_half((uint16_t)(pk.z >> 16)));
    v[3] = __half2float(__ushort_as_half((uint16_t)(pk.w >> 16)));
  }
};
template <> struct BlkEnt<ValF32> {
  static __device__ __forceinline__ void load(const BlkArgs& a, long long e, unsigned cm,
                                              unsigned c[4], float v[4]) {
    typedef unsigned short us4 __attribute__((ext_vector_type(4)));
    const us4 cc = *reinterpret_cast<const us4*>(static_cast<const uint16_t*>(a.ent) + e);
    const float4 vv = *reinterpret_cast<const float4*>(a.eval + e);
    c[0] = cc.x & cm; c[1] = cc.y & cm; c[2] = cc.z & cm; c[3] = cc.w & cm;
    v[0] = vv.x; v[1] = vv.y; v[2] = vv.z; v[3] = vv.w;
  }
};

// ------------------------------------------------------------------ pieces
// Tile h of G into LDS: xs[v Fp + f] = G[colinv[h C + v] F + f] (0 for padding
// positions and frames past F).
__device__ __forceinline__ void blk_stage(const BlkArgs& a, int h, float* xs) {
  const int fq = a.Fp >> 2;
  const int lgfq = a.lgFp - 2;
  const int n4 = a.C << lgfq;
  const int* ci = a.colinv + (long long)h * a.C;
  for (int i = threadIdx.x; i < n4; i += BLK_THREADS) {
    const int v = i >> lgfq, f = (i & (fq - 1)) << 2;
    const int c = ci[v];
    float4 val = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (c >= 0 && f < a.F) {
      const float* src = a.G + (long long)c * a.F + f;
      if (a.F4) {
        val = *reinterpret_cast<const float4*>(src);
      } else {
        val.x = src[0];
        if (f + 1 < a.F) val.y = src[1];
        if (f + 2 < a.F) val.z = src[2];
        if (f + 3 < a.F) val.w = src[3];
      }
    }
    reinterpret_cast<float4*>(xs)[i] = val;
  }
}

// Segment starts of the block's rbe row positions in tile h, relative to the
// block's first entry in the tile: s_start[r] (exclusive), s_start[rbe] = total.
// Part 1 (before a barrier) scans within wavefronts; part 2 adds the wavefront
// bases.
__device__ __forceinline__ void blk_starts1(const BlkArgs& a, int h, int p0, int rbe,
                                            int* s_start, int* s_wsum, int loc[4], int* tot) {
  const int t = threadIdx.x;
  const int r = 4 * t;
  const uint16_t* cn = a.cnt + (long long)h * a.m_pad + p0;
  int v[4] = {0, 0, 0, 0};
  if (r < rbe) {
    typedef unsigned short us4 __attribute__((ext_vector_type(4)));
    const us4 c4 = *reinterpret_cast<const us4*>(cn + r);   // rbe is a multiple of 64
    v[0] = c4.x; v[1] = c4.y; v[2] = c4.z; v[3] = c4.w;
  }
  loc[0] = 0;
  loc[1] = v[0];
  loc[2] = v[0] + v[1];
  loc[3] = v[0] + v[1] + v[2];
  const int mine = loc[3] + v[3];
  int x = mine;
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const int y = __shfl_up(x, off, WAVE);
    if (lane >= off) x += y;
  }
  if (lane == WAVE - 1) s_wsum[wave_id()] = x;
  *tot = x - mine;   // exclusive within the wavefront
}

__device__ __forceinline__ void blk_starts2(int rbe, int* s_start, const int* s_wsum,
                                            const int loc[4], int excl) {
  const int t = threadIdx.x;
  int base = excl;
  const int w = wave_id();
  int all = 0;
#pragma unroll
  for (int k = 0; k < BLK_WAVES; k++) {
    const int c = s_wsum[k];
    base += k < w ? c : 0;
    all += c;
  }
  const int r = 4 * t;
  if (r < rbe) {
    s_start[r] = base + loc[0];
    s_start[r + 1] = base + loc[1];
    s_start[r + 2] = base + loc[2];
    s_start[r + 3] = base + loc[3];
  }
  if (t == 0) s_start[rbe] = all;
}

// Tile range [h0, h1) of split sp of a block: tiles whose entry prefix (over
// the block's chunks, in tile order) falls in [tot sp / splits, tot (sp+1) /
// splits).  Wavefront 0 computes it; the caller holds barriers around it.
__device__ __forceinline__ void blk_split_range(const BlkArgs& a, int j0, int j1, int sp,
                                                int* s_range) {
  if (wave_id() != 0) return;
  const int lane = lane_id();
  const int per = (a.T + WAVE - 1) / WAVE;
  const int hb = min(a.T, lane * per), he = min(a.T, hb + per);
  long long mine = 0;
  for (int h = hb; h < he; h++) {
    const int64_t* sb = a.seg_base + (long long)h * (a.nG + 1);
    mine += sb[j1] - sb[j0];
  }
  long long x = mine;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const long long y = __shfl_up(x, off, WAVE);
    if (lane >= off) x += y;
  }
  const long long tot = __shfl(x, WAVE - 1, WAVE);
  const long long lo = tot * sp / a.splits, hi = tot * (sp + 1) / a.splits;
  if (lane == 0) { s_range[0] = a.T; s_range[1] = a.T; }
  __builtin_amdgcn_wave_barrier();
  long long pre = x - mine;   // entries before tile hb
  for (int h = hb; h < he; h++) {
    const int64_t* sb = a.seg_base + (long long)h * (a.nG + 1);
    // the prefix is nondecreasing: the smallest tile whose prefix reaches lo
    // (resp. hi) starts the split (resp. the next one)
    if (pre >= lo) atomicMin(&s_range[0], h);
    if (pre >= hi) atomicMin(&s_range[1], h);
    pre += sb[j1] - sb[j0];
  }
  __builtin_amdgcn_wave_barrier();
  if (lane == 0 && sp == a.splits - 1) s_range[1] = a.T;
  if (lane == 0 && sp == 0) s_range[0] = 0;
}

// ------------------------------------------------------------------ main kernel
template <typename VT, int P, int K>
__global__ __launch_bounds__(BLK_THREADS) void k_blk(BlkArgs a) {
  extern __shared__ float4 blk_lds[];
  float* xs = reinterpret_cast<float*>(blk_lds);
  int* s_start = reinterpret_cast<int*>(xs + ((size_t)a.C << a.lgFp));
  __shared__ int s_wsum[BLK_WAVES];
  __shared__ int s_range[2];
  __shared__ double s_ll[64];
  const int lane = lane_id();
  const int w = wave_id();
  const int slot = lane / a.L, q = lane % a.L;
  const unsigned cm = (unsigned)a.C - 1u;
  if (threadIdx.x < 64) s_ll[threadIdx.x] = 0.0;
  const long long nwork = (long long)a.nblocks * a.splits;
  for (long long work = blockIdx.x; work < nwork; work += gridDim.x) {
    const int blk = (int)(work / a.splits), sp = (int)(work % a.splits);
    const int p0 = blk * a.rb;
    const int rbe = min(a.rb, a.m_pad - p0);
    const int j0 = p0 / BLK_G, j1 = (p0 + rbe) / BLK_G;
    int h0 = 0, h1 = a.T;
    if (a.splits > 1) {
      __syncthreads();
      blk_split_range(a, j0, j1, sp, s_range);
      __syncthreads();
      h0 = s_range[0];
      h1 = s_range[1];
    }
    float acc[K][P];
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
      for (int j = 0; j < P; j++) acc[k][j] = 0.0f;
    for (int h = h0; h < h1; h++) {
      const int64_t* sb = a.seg_base + (long long)h * (a.nG + 1);
      const long long c0 = sb[j0], c1 = sb[j1];
      if (c0 == c1) continue;   // the block has no entry in this tile
      __syncthreads();          // the previous tile's readers are done
      blk_stage(a, h, xs);
      int loc[4], excl;
      blk_starts1(a, h, p0, rbe, s_start, s_wsum, loc, &excl);
      __syncthreads();
      blk_starts2(rbe, s_start, s_wsum, loc, excl);
      __syncthreads();
      const float* xq = xs + q * P;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const int r = (w * K + k) * a.S + slot;
        long long e = 0, en = 0;
        if (r < rbe) {
          e = c0 + s_start[r];
          en = c0 + s_start[r + 1];
        }
        for (;;) {
          const bool act = e < en;
          if (!__any(act)) break;
          if (act) {
            // two 4-entry groups in flight; the second is zeroed past the segment
            const bool two = e + 4 < en;
            unsigned c[8];
            float v[8];
            BlkEnt<VT>::load(a, e, cm, c, v);
            BlkEnt<VT>::load(a, two ? e + 4 : e, cm, c + 4, v + 4);
#pragma unroll
            for (int t = 4; t < 8; t++) v[t] = two ? v[t] : 0.0f;
#pragma unroll
            for (int t = 0; t < 8; t++) {
              const float4* xp = reinterpret_cast<const float4*>(xq + (c[t] << a.lgFp));
#pragma unroll
              for (int i = 0; i < P / 4; i++) {
                const float4 xv = xp[i];
                acc[k][4 * i + 0] += v[t] * xv.x;
                acc[k][4 * i + 1] += v[t] * xv.y;
                acc[k][4 * i + 2] += v[t] * xv.z;
                acc[k][4 * i + 3] += v[t] * xv.w;
              }
            }
          }
          e += 8;
        }
      }
    }
    // ---- epilogue: the block's rows are complete (or this split's share is)
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int r = (w * K + k) * a.S + slot;
      if (r >= rbe) continue;
      const int p = p0 + r;
      if (a.mode == 2) {
        float* dst = a.partial + (((long long)sp * a.m_pad + p) << a.lgFp) + q * P;
#pragma unroll
        for (int i = 0; i < P / 4; i++)
          reinterpret_cast<float4*>(dst)[i] =
              make_float4(acc[k][4 * i], acc[k][4 * i + 1], acc[k][4 * i + 2], acc[k][4 * i + 3]);
        continue;
      }
      const int orig = a.rowperm[p];
      if (orig < 0) continue;
      if (a.mode == 1) {
#pragma unroll
        for (int j = 0; j < P; j++) {
          const int f = q * P + j;
          if (f < a.F) a.out[(long long)orig * a.F + f] = acc[k][j];
        }
      } else {
        const float sc = a.rowscale ? a.rowscale[orig] : 1.0f;
#pragma unroll
        for (int j = 0; j < P; j++) {
          const int f = q * P + j;
          if (f >= a.F) continue;
          const long long o = (long long)orig * a.F + f;
          const float yh = acc[k][j] * sc;
          const float yi = a.Y[o];
          const float ev = yh > 0.0f ? yi / yh : 0.0f;
          a.out[o] = a.fold ? ev * sc : ev;
          if (yh > 0.0f && a.loglik) atomicAdd(&s_ll[f], (double)yi * log((double)yh) - (double)yh);
        }
      }
    }
  }
  if (a.mode == 0 && a.loglik) {
    __syncthreads();
    if (threadIdx.x < a.F) unsafeAtomicAdd(&a.loglik[(long long)(*a.itp) * a.F + threadIdx.x],
                                           s_ll[threadIdx.x]);
  }
}

// Split reductions, in split order.  Forward: the MLEM ratio and the FP64
// log-likelihood (one atomic per workgroup and frame); back projection: the image.
__global__ __launch_bounds__(256) void k_blk_reduce(BlkArgs a) {
  __shared__ double s_ll[64];
  if (threadIdx.x < 64) s_ll[threadIdx.x] = 0.0;
  __syncthreads();
  const long long tot = (long long)a.m_pad * a.F;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < tot;
       i += (long long)gridDim.x * blockDim.x) {
    const long long p = i / a.F;
    const int f = (int)(i % a.F);
    const int orig = a.rowperm[p];
    if (orig < 0) continue;
    float s = 0.0f;
    for (int sp = 0; sp < a.splits; sp++)
      s += a.partial[(((long long)sp * a.m_pad + p) << a.lgFp) + f];
    const long long o = (long long)orig * a.F + f;
    if (a.mode == 1) {
      a.out[o] = s;
    } else {
      const float sc = a.rowscale ? a.rowscale[orig] : 1.0f;
      const float yh = s * sc;
      const float yi = a.Y[o];
      const float ev = yh > 0.0f ? yi / yh : 0.0f;
      a.out[o] = a.fold ? ev * sc : ev;
      if (yh > 0.0f && a.loglik) atomicAdd(&s_ll[f], (double)yi * log((double)yh) - (double)yh);
    }
  }
  if (a.mode == 0 && a.loglik) {
    __syncthreads();
    if (threadIdx.x < a.F) unsafeAtomicAdd(&a.loglik[(long long)(*a.itp) * a.F + threadIdx.x],
                                           s_ll[threadIdx.x]);
  }
}

// ------------------------------------------------------------------ build kernels
// Per row position: entries per tile (LDS histogram of T bins per wavefront),
// written padded to a multiple of 4 into cnt32[h m_pad + p] (zero-filled before).
__global__ __launch_bounds__(256) void k_blk_count(int m_pad, int T, int lgC,
                                                   const int* __restrict__ rowperm,
                                                   const int64_t* __restrict__ rowptr,
                                                   const int* __restrict__ col,
                                                   const int* __restrict__ cpos,
                                                   int* __restrict__ cnt32) {
  extern __shared__ int bins[];   // (256 / WAVE) x T
  const int lane = lane_id();
  int* hb = bins + wave_id() * T;
  const long long nw = (long long)gridDim.x * (256 / WAVE);
  for (long long p = (long long)blockIdx.x * (256 / WAVE) + wave_id(); p < m_pad; p += nw) {
    const int r = rowperm[p];
    if (r < 0) continue;
    for (int i = lane; i < T; i += WAVE) hb[i] = 0;
    __builtin_amdgcn_wave_barrier();
    const long long b = rowptr[r], e = rowptr[r + 1];
    for (long long k = b + lane; k < e; k += WAVE) atomicAdd(&hb[cpos[col[k]] >> lgC], 1);
    __builtin_amdgcn_wave_barrier();
    for (int i = lane; i < T; i += WAVE) {
      const int c = hb[i];
      if (c) cnt32[(long long)i * m_pad + p] = (c + 3) & ~3;
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Per row position: entries placed in their (tile, row) segments in the row's
// column order (stable: ranks by ballot within 64-entry chunks), then zero
// entries up to the padded length.
template <typename V>
__global__ __launch_bounds__(256) void k_blk_fill(int m_pad, int T, int lgC,
                                                  const int* __restrict__ rowperm,
                                                  const int64_t* __restrict__ rowptr,
                                                  const int* __restrict__ col,
                                                  const V* __restrict__ val,
                                                  const int* __restrict__ cpos,
                                                  const int64_t* __restrict__ off,
                                                  void* __restrict__ ent,
                                                  float* __restrict__ eval) {
  extern __shared__ int bins[];   // (256 / WAVE) x T running counts
  const int lane = lane_id();
  int* hb = bins + wave_id() * T;
  const unsigned cm = (1u << lgC) - 1u;
  const long long nw = (long long)gridDim.x * (256 / WAVE);
  for (long long p = (long long)blockIdx.x * (256 / WAVE) + wave_id(); p < m_pad; p += nw) {
    const int r = rowperm[p];
    if (r < 0) continue;
    for (int i = lane; i < T; i += WAVE) hb[i] = 0;
    __builtin_amdgcn_wave_barrier();
    const long long b = rowptr[r], e = rowptr[r + 1];
    for (long long base = b; base < e; base += WAVE) {
      const long long k = base + lane;
      bool act = k < e;
      const int pos = act ? cpos[col[k]] : 0;
      const int h = act ? (pos >> lgC) : -1;
      unsigned long long left = __ballot(act);
      while (left) {
        const int lead = __ffsll((long long)left) - 1;
        const int key = __shfl(h, lead);
        const unsigned long long mask = __ballot(act && h == key);
        const int start = hb[key];
        if (act && h == key) {
          const int rank = __popcll(mask & ((1ull << lane) - 1ull));
          const long long dst = off[(long long)key * m_pad + p] + start + rank;
          const unsigned local = (unsigned)pos & cm;
          if (sizeof(V) == 2) {
            static_cast<uint32_t*>(ent)[dst] = local | ((uint32_t)(uint16_t)val[k] << 16);
          } else {
            static_cast<uint16_t*>(ent)[dst] = (uint16_t)local;
            eval[dst] = (float)val[k];
          }
          act = false;
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == lead) hb[key] = start + __popcll(mask);
        __builtin_amdgcn_wave_barrier();
        left &= ~mask;
      }
    }
    __builtin_amdgcn_wave_barrier();
    // zero entries up to the padded segment length
    for (int i = lane; i < T; i += WAVE) {
      const int c = hb[i];
      const int cp = (c + 3) & ~3;
      const long long d0 = off[(long long)i * m_pad + p];
      for (int t = c; t < cp; t++) {
        if (sizeof(V) == 2) {
          static_cast<uint32_t*>(ent)[d0 + t] = 0u;
        } else {
          static_cast<uint16_t*>(ent)[d0 + t] = 0;
          eval[d0 + t] = 0.0f;
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

__global__ __launch_bounds__(256) void k_blk_narrow(const int* __restrict__ c32, long long n,
                                                    uint16_t* __restrict__ c16) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c16[i] = (uint16_t)c32[i];
}

// seg_base[h (nG+1) + j] = off[h m_pad + j BLK_G] (j = nG: the tile's end)
__global__ __launch_bounds__(256) void k_blk_bases(int T, int nG, int m_pad,
                                                   const int64_t* __restrict__ off,
                                                   int64_t* __restrict__ sb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)T * (nG + 1)) return;
  const long long h = i / (nG + 1), j = i % (nG + 1);
  sb[i] = off[h * m_pad + j * BLK_G];
}

__global__ __launch_bounds__(256) void k_fill_i32_v(int* __restrict__ p, long long n, int v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_set_i64_at(int64_t* p, long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// ------------------------------------------------------------------ orders (host)
static int ilog2(int v) {
  int l = 0;
  while ((1 << l) < v) l++;
  return l;
}

static uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  uint32_t r = 0;
  for (int b = 0; b < 10; b++)
    r |= (((x >> b) & 1u) << (3 * b + 2)) | (((y >> b) & 1u) << (3 * b + 1)) |
         (((z >> b) & 1u) << (3 * b));
  return r;
}
static uint32_t morton2(uint32_t x, uint32_t y) {
  uint32_t r = 0;
  for (int b = 0; b < 16; b++) r |= (((x >> b) & 1u) << (2 * b + 1)) | (((y >> b) & 1u) << (2 * b));
  return r;
}

// Voxel order: 8x8x8 bricks in Morton order of their brick coordinates, voxels
// inside a brick in id order.  Identity when the grid is unknown.
static std::vector<int> voxel_order(const MatGeo& g, int n) {
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  if ((long long)g.nx * g.ny * g.nz != n || n == 0) return ord;
  std::vector<std::pair<uint64_t, int>> key(n);
  for (int v = 0; v < n; v++) {
    const int iz = v % g.nz, iy = (v / g.nz) % g.ny, ix = v / (g.nz * g.ny);
    key[v] = {((uint64_t)morton3(ix >> 3, iy >> 3, iz >> 3) << 32) | (uint32_t)v, v};
  }
  std::sort(key.begin(), key.end());
  for (int i = 0; i < n; i++) ord[i] = key[i].second;
  return ord;
}

// Channel order: by angle, then by the line's midpoint on the detector plane
// (id0y + id1y, id0z + id1z) in Morton order, then by its slope: lines that
// cross the same voxels come together.  Identity without channel ids.
static std::vector<int> channel_order(const MatGeo& g, const std::vector<int>& ids) {
  const int m = (int)ids.size();
  std::vector<int> ord(m);
  std::iota(ord.begin(), ord.end(), 0);
  if (g.n1y == 0 || m == 0) return ord;
  const long long P1 = g.n1y, P2 = P1 * g.n1z, P3 = P2 * g.n0y, P4 = P3 * g.n0z;
  std::vector<std::pair<uint64_t, int>> key(m);
  for (int i = 0; i < m; i++) {
    const long long c = ids[i];
    const long long ida = c / P4, t = c % P4;
    const int id0z = (int)(t / P3), id0y = (int)((t % P3) / P2);
    const int id1z = (int)((t % P2) / P1), id1y = (int)(t % P1);
    const uint32_t my = (uint32_t)(id0y + id1y), mz = (uint32_t)(id0z + id1z);
    const uint32_t dy = (uint32_t)(id1y - id0y + 64), dz = (uint32_t)(id1z - id0z + 64);
    const uint64_t k = ((uint64_t)ida << 48) | ((uint64_t)morton2(mz, my) << 16) |
                       ((uint64_t)(dz & 0xFF) << 8) | (uint64_t)(dy & 0xFF);
    key[i] = {k, i};
  }
  std::stable_sort(key.begin(), key.end(),
                   [](const std::pair<uint64_t, int>& x, const std::pair<uint64_t, int>& y) {
                     return x.first < y.first;
                   });
  for (int i = 0; i < m; i++) ord[i] = key[i].second;
  return ord;
}

// ------------------------------------------------------------------ build (host)
// Blocked copy of M (rows, columns) with row order `rows` (positions ->
// rows of M) and column order `cols` (positions -> columns of M).  Returns
// EPET_OK with *out == nullptr when memory is short.
static int blk_build(epet_ctx* ctx, const epet_matrix* M, int Fp, const std::vector<int>& rows,
                     const std::vector<int>& cols, BlkCopy** out) {
  *out = nullptr;
  hipStream_t st = ctx->stream;
  BlkCopy* B = new BlkCopy();
  B->Fp = Fp;
  B->lgFp = ilog2(Fp);
  B->C = BLK_TILE_FLOATS / Fp;
  B->lgC = ilog2(B->C);
  B->m = M->m;
  B->m_pad = (int)(((long long)M->m + BLK_G - 1) / BLK_G * BLK_G);
  B->nG = B->m_pad / BLK_G;
  B->T = (int)(((long long)M->n + B->C - 1) / B->C);
  B->vt = M->value_type;
  const long long nseg = (long long)B->T * B->m_pad;
  std::vector<int> rp(B->m_pad, -1), ci((size_t)B->T * B->C, -1), cpos(std::max(1, M->n), 0);
  for (int i = 0; i < M->m; i++) rp[i] = rows[i];
  for (int i = 0; i < M->n; i++) {
    ci[i] = cols[i];
    cpos[cols[i]] = i;
  }
  int* cpos_d = nullptr;
  int* cnt32 = nullptr;
  int64_t* off = nullptr;
  bool ok = hipMalloc((void**)&B->rowperm, sizeof(int) * (size_t)B->m_pad) == hipSuccess &&
            hipMalloc((void**)&B->colinv, sizeof(int) * ci.size()) == hipSuccess &&
            hipMalloc((void**)&cpos_d, sizeof(int) * cpos.size()) == hipSuccess &&
            hipMalloc((void**)&B->seg_base, sizeof(int64_t) * (size_t)B->T * (B->nG + 1)) ==
                hipSuccess &&
            hipMalloc((void**)&B->cnt, sizeof(uint16_t) * (size_t)nseg) == hipSuccess &&
            hipMalloc((void**)&cnt32, sizeof(int) * (size_t)nseg) == hipSuccess &&
            hipMalloc((void**)&off, sizeof(int64_t) * (size_t)(nseg + 1)) == hipSuccess;
  auto cleanup = [&]() {
    if (cpos_d) (void)hipFree(cpos_d);
    if (cnt32) (void)hipFree(cnt32);
    if (off) (void)hipFree(off);
  };
  if (!ok) {
    (void)hipGetLastError();
    cleanup();
    delete B;
    return EPET_OK;
  }
  EPET_HIP(hipMemcpyAsync(B->rowperm, rp.data(), sizeof(int) * rp.size(), hipMemcpyHostToDevice, st));
  EPET_HIP(hipMemcpyAsync(B->colinv, ci.data(), sizeof(int) * ci.size(), hipMemcpyHostToDevice, st));
  EPET_HIP(hipMemcpyAsync(cpos_d, cpos.data(), sizeof(int) * cpos.size(), hipMemcpyHostToDevice, st));
  k_fill_i32_v<<<grid_for(nseg, 256, 1ll << 31), 256, 0, st>>>(cnt32, nseg, 0);
  const size_t lds = sizeof(int) * (256 / WAVE) * (size_t)B->T;
  int rc = EPET_OK;
  if (lds > 160 * 1024) {
    rc = fail(EPET_ERR_UNSUPPORTED, __FILE__, __LINE__, "blocked copy: too many column tiles");
  }
  const unsigned g = (unsigned)std::max<long long>(
      1, std::min<long long>(((long long)B->m_pad + 3) / 4, 16ll * ctx->num_cus));
  if (rc == EPET_OK) {
    if (lds > 64 * 1024) {
      (void)hipFuncSetAttribute((const void*)k_blk_count, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)lds);
      (void)hipFuncSetAttribute((const void*)k_blk_fill<float>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      (void)hipFuncSetAttribute((const void*)k_blk_fill<uint16_t>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    }
    k_blk_count<<<g, 256, lds, st>>>(B->m_pad, B->T, B->lgC, B->rowperm, M->rowptr, M->col, cpos_d,
                                     cnt32);
    if (hipGetLastError() != hipSuccess) rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "k_blk_count");
  }
  long long tot = 0;
  if (rc == EPET_OK) rc = exclusive_scan_i32_to_i64(ctx, cnt32, off, nseg, &tot);
  if (rc == EPET_OK) {
    B->nent = tot;
    k_set_i64_at<<<1, 64, 0, st>>>(off + nseg, tot);
    k_blk_narrow<<<grid_for(nseg, 256, 1ll << 31), 256, 0, st>>>(cnt32, nseg, B->cnt);
    k_blk_bases<<<grid_for((long long)B->T * (B->nG + 1), 256, 1ll << 31), 256, 0, st>>>(
        B->T, B->nG, B->m_pad, off, B->seg_base);
    (void)hipFree(cnt32);
    cnt32 = nullptr;
    const size_t eb = M->value_type == EPET_VAL_F32 ? 2 : 4;
    if (hipMalloc(&B->ent, eb * (size_t)std::max(4ll, tot)) != hipSuccess ||
        (M->value_type == EPET_VAL_F32 &&
         hipMalloc((void**)&B->eval, sizeof(float) * (size_t)std::max(4ll, tot)) != hipSuccess)) {
      (void)hipGetLastError();
      cleanup();
      delete B;
      return EPET_OK;
    }
    if (M->value_type == EPET_VAL_F32)
      k_blk_fill<float><<<g, 256, lds, st>>>(B->m_pad, B->T, B->lgC, B->rowperm, M->rowptr, M->col,
                                             (const float*)M->val, cpos_d, off, B->ent, B->eval);
    else
      k_blk_fill<uint16_t><<<g, 256, lds, st>>>(B->m_pad, B->T, B->lgC, B->rowperm, M->rowptr,
                                                M->col, (const uint16_t*)M->val, cpos_d, off,
                                                B->ent, nullptr);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
      rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "blocked copy fill");
  }
  cleanup();
  if (rc != EPET_OK) {
    delete B;
    return rc;
  }
  *out = B;
  return EPET_OK;
}

// ------------------------------------------------------------------ launch
typedef void (*BlkKern)(BlkArgs);

template <typename VT>
static BlkKern blk_kernel(int P, int K) {
#define EPET_BK(PP, KK) \
  if (P == PP && K == KK) return k_blk<VT, PP, KK>;
  EPET_BK(4, 1) EPET_BK(4, 2) EPET_BK(4, 4) EPET_BK(4, 8)
  EPET_BK(8, 1) EPET_BK(8, 2) EPET_BK(8, 4)
  EPET_BK(16, 1) EPET_BK(16, 2)
  EPET_BK(32, 1)
#undef EPET_BK
  return nullptr;
}

static BlkKern blk_kernel_vt(int vt, int P, int K) {
  switch (vt) {
    case EPET_VAL_F32: return blk_kernel<ValF32>(P, K);
    case EPET_VAL_BF16: return blk_kernel<ValBF16>(P, K);
    default: return blk_kernel<ValF16>(P, K);
  }
}

// rows per block of a (P, K) shape for Fp padded frames
static int blk_rb(int Fp, int P, int K) {
  const int L = Fp / P;
  return BLK_WAVES * (WAVE / L) * K;
}

// K P accumulators per lane: at most 32, so that 1024-thread workgroups keep
// within 128 VGPRs without spilling
static bool blk_shape_valid(int Fp, int P, int K) {
  if (P > Fp || K * P > 32) return false;
  const int L = Fp / P;
  if (L > WAVE) return false;
  const int rb = blk_rb(Fp, P, K);
  return rb % BLK_G == 0 && rb <= BLK_MAX_RB;
}

// work-item splits of a direction: enough work items for every CU twice over
static int blk_splits(const BlkCopy* B, int rb, int ncu) {
  const int nblocks = (B->m_pad + rb - 1) / rb;
  const int want = 2 * ncu;
  if (nblocks >= want) return 1;
  return std::min(B->T, std::max(1, (want + nblocks - 1) / nblocks));
}

static int blk_partial_ensure(BlkCopy* B, int splits) {
  if (splits <= 1) return EPET_OK;
  const long long need = (long long)splits * B->m_pad * B->Fp;
  if (need <= B->partial_len) return EPET_OK;
  if (B->partial) (void)hipFree(B->partial);
  B->partial = nullptr;
  B->partial_len = 0;
  if (hipMalloc((void**)&B->partial, sizeof(float) * (size_t)need) != hipSuccess) {
    (void)hipGetLastError();
    EPET_FAIL(EPET_ERR_ALLOC, "blocked products: partial sums");
  }
  B->partial_len = need;
  return EPET_OK;
}

// One product with an explicit shape (mode 0 forward MLEM step, 1 image).
static int blk_launch(epet_ctx* ctx, const BlkCopy* B, int P, int K, int splits, int F,
                      const float* G, float* out, int mode, const float* rowscale, int fold,
                      const float* Y, const int* itp, double* loglik) {
  BlkKern kern = blk_kernel_vt(B->vt, P, K);
  EPET_CHECK(kern, EPET_ERR_STATE, "blocked products: no kernel for this shape");
  BlkArgs a{};
  a.F = F;
  a.Fp = B->Fp;
  a.lgFp = B->lgFp;
  a.C = B->C;
  a.T = B->T;
  a.m_pad = B->m_pad;
  a.nG = B->nG;
  a.rb = blk_rb(B->Fp, P, K);
  a.nblocks = (B->m_pad + a.rb - 1) / a.rb;
  a.splits = splits;
  a.L = B->Fp / P;
  a.S = WAVE / a.L;
  a.mode = splits > 1 ? 2 : mode;
  a.F4 = (F % 4 == 0) ? 1 : 0;
  a.rowperm = B->rowperm;
  a.colinv = B->colinv;
  a.seg_base = B->seg_base;
  a.cnt = B->cnt;
  a.ent = B->ent;
  a.eval = B->eval;
  a.G = G;
  a.out = out;
  a.partial = B->partial;
  a.rowscale = rowscale;
  a.fold = fold;
  a.Y = Y;
  a.itp = itp;
  a.loglik = loglik;
  EPET_CHECK(splits <= 1 || (B->partial && B->partial_len >= (long long)splits * B->m_pad * B->Fp),
             EPET_ERR_STATE, "blocked products: partial buffer missing");
  const size_t lds = sizeof(float) * ((size_t)B->C << B->lgFp) + sizeof(int) * (BLK_MAX_RB + 4);
  static bool attr_set[2][3][64] = {};
  const int pi = ilog2(P), ki = ilog2(K);
  bool& done = attr_set[0][B->vt][pi * 8 + ki];
  if (!done) {
    (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    done = true;
  }
  const long long nwork = (long long)a.nblocks * a.splits;
  const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(nwork, ctx->num_cus));
  kern<<<grid, BLK_THREADS, lds, ctx->stream>>>(a);
  if (splits > 1) {
    a.mode = mode;
    k_blk_reduce<<<grid_for((long long)B->m_pad * F, 256, 16ll * ctx->num_cus), 256, 0,
                   ctx->stream>>>(a);
  }
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

static int padded_frames(int F) {
  int Fp = 8;
  while (Fp < F) Fp <<= 1;
  return Fp;
}

static BlkCopy* blk_find(const epet_matrix* M, int Fp) {
  for (BlkCopy* b = M->blk; b; b = b->next)
    if (b->Fp == Fp) return b;
  return nullptr;
}

// Builds the blocked copies of A (forward) and of its transposed copy (back
// projection) for F frames, once per padded frame count (allocates and
// synchronises: call outside graph capture).  Without memory nothing is built
// and the gather kernels run instead.
int blk_prepare(epet_ctx* ctx, epet_matrix* A, int F) {
  epet_matrix* At = A->tr;
  if (!At || F < 2 || F > 64 || !A->cols_strict || A->m == 0 || A->n == 0) return EPET_OK;
  const int Fp = padded_frames(F);
  if (blk_find(A, Fp) && blk_find(At, Fp)) return EPET_OK;
  // orders: channel rows by line, voxels by brick
  std::vector<int> ids(A->m);
  if (A->row_ids)
    EPET_HIP(hipMemcpy(ids.data(), A->row_ids, sizeof(int) * A->m, hipMemcpyDeviceToHost));
  const std::vector<int> ch = A->row_ids ? channel_order(A->mg, ids) : channel_order(MatGeo{}, ids);
  const std::vector<int> vx = voxel_order(A->mg, A->n);
  if (!blk_find(A, Fp)) {
    BlkCopy* B = nullptr;
    EPET_TRY(blk_build(ctx, A, Fp, ch, vx, &B));
    if (!B) return EPET_OK;
    B->next = A->blk;
    A->blk = B;
  }
  if (!blk_find(At, Fp)) {
    BlkCopy* B = nullptr;
    EPET_TRY(blk_build(ctx, At, Fp, vx, ch, &B));
    if (!B) return EPET_OK;
    B->next = At->blk;
    At->blk = B;
  }
  return EPET_OK;
}

bool blk_built(const epet_matrix* A, int F) {
  if (!A->tr || F < 2 || F > 64) return false;
  const int Fp = padded_frames(F);
  return blk_find(A, Fp) && blk_find(A->tr, Fp);
}

bool blk_ready(const epet_matrix* A, int F) {
  if (!A->tr || F < 2 || F > 64) return false;
  const int Fp = padded_frames(F);
  const BlkCopy* f = blk_find(A, Fp);
  const BlkCopy* b = blk_find(A->tr, Fp);
  if (!f || !b || f->tuned_F != F || b->tuned_F != F) return false;
  auto ok = [](const BlkCopy* B) {
    return B->splits <= 1 || B->partial_len >= (long long)B->splits * B->m_pad * B->Fp;
  };
  return ok(f) && ok(b);
}

// Forward + back projection of one multi-frame MLEM iteration with the tuned
// shapes (capturable once prepared and tuned).
int blk_iteration(epet_ctx* ctx, const epet_matrix* A, const MfArgs& a) {
  const int Fp = padded_frames(a.F);
  const BlkCopy* f = blk_find(A, Fp);
  const BlkCopy* b = blk_find(A->tr, Fp);
  EPET_CHECK(f && b, EPET_ERR_STATE, "blocked copies missing");
  EPET_TRY(blk_launch(ctx, f, f->P, f->K, f->splits, a.F, a.x, a.e, 0, A->rowscale,
                      A->rowscale ? 1 : 0, a.y, a.itp, a.loglik));
  EPET_TRY(blk_launch(ctx, b, b->P, b->K, b->splits, a.F, a.e, a.c, 1, nullptr, 0, nullptr,
                      nullptr, nullptr));
  return EPET_OK;
}

// Autotuner: every valid (frames per lane P, row groups per lane K) shape of
// each direction is timed with HIP events on the context's stream (reps
// launches after one warm-up) and the fastest kept.
int blk_tune(epet_ctx* ctx, epet_matrix* A, const MfArgs& a, int reps) {
  const int Fp = padded_frames(a.F);
  BlkCopy* dirs[2] = {blk_find(A, Fp), blk_find(A->tr, Fp)};
  EPET_CHECK(dirs[0] && dirs[1], EPET_ERR_STATE, "blocked copies missing");
  hipEvent_t e0, e1;
  EPET_HIP(hipEventCreate(&e0));
  EPET_HIP(hipEventCreate(&e1));
  static const int Ps[4] = {4, 8, 16, 32};
  static const int Ks[4] = {1, 2, 4, 8};
  int rc = EPET_OK;
  for (int d = 0; d < 2 && rc == EPET_OK; d++) {
    BlkCopy* B = dirs[d];
    float best = 1e30f;
    int bP = 0, bK = 0, bS = 1;
    for (int pi = 0; pi < 4 && rc == EPET_OK; pi++)
      for (int ki = 0; ki < 4 && rc == EPET_OK; ki++) {
        const int P = Ps[pi], K = Ks[ki];
        if (!blk_shape_valid(Fp, P, K)) continue;
        const int splits = blk_splits(B, blk_rb(Fp, P, K), ctx->num_cus);
        if (blk_partial_ensure(B, splits) != EPET_OK) { (void)hipGetLastError(); continue; }
        auto run = [&]() {
          return d == 0 ? blk_launch(ctx, B, P, K, splits, a.F, a.x, a.e, 0, A->rowscale,
                                     A->rowscale ? 1 : 0, a.y, a.itp, nullptr)
                        : blk_launch(ctx, B, P, K, splits, a.F, a.e, a.c, 1, nullptr, 0, nullptr,
                                     nullptr, nullptr);
        };
        rc = run();
        if (rc != EPET_OK) break;
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) { rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "event"); break; }
        for (int r = 0; r < reps && rc == EPET_OK; r++) rc = run();
        if (rc != EPET_OK) break;
        float ms = 0.0f;
        if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "tuning events");
          break;
        }
        ms /= reps;
        if (ms < best) { best = ms; bP = P; bK = K; bS = splits; }
      }
    if (rc == EPET_OK) {
      EPET_CHECK(bP > 0, EPET_ERR_STATE, "blocked products: no valid shape");
      B->P = bP;
      B->K = bK;
      B->splits = bS;
      B->ms = best;
      B->tuned_F = a.F;
      rc = blk_partial_ensure(B, bS);
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// Launch shape of the blocked products for F frames: {1, P_fwd, K_fwd,
// rows per block fwd, splits fwd, P_bwd, K_bwd, rows per block bwd, splits bwd}
// and the tuned times, or {0} when the blocked path is not in use.
int blk_info(const epet_matrix* A, int F, int info[9], float ms[2]) {
  for (int i = 0; i < 9; i++) info[i] = 0;
  ms[0] = ms[1] = 0.0f;
  if (!blk_ready(A, F)) return EPET_OK;
  const int Fp = padded_frames(F);
  const BlkCopy* f = blk_find(A, Fp);
  const BlkCopy* b = blk_find(A->tr, Fp);
  info[0] = 1;
  info[1] = f->P; info[2] = f->K; info[3] = blk_rb(Fp, f->P, f->K); info[4] = f->splits;
  info[5] = b->P; info[6] = b->K; info[7] = blk_rb(Fp, b->P, b->K); info[8] = b->splits;
  ms[0] = f->ms;
  ms[1] = b->ms;
  return EPET_OK;
}

}  // namespace epet
