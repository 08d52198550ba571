// blk.hip -- batched (multi-frame) projections for MLEM with the operator
// streamed once per iteration and direction.
//
// F frames share one matrix M (m rows, n columns); the gathered vector G and
// the output are frame-minor (G[c F + f]).  The matrix is kept once more as a
// blocked copy:
//   * rows in a locality order (rowperm: position -> row) and columns in a
//     locality order (colinv: position -> column), both cut into pieces: row
//     blocks of RB positions (chosen at launch, a multiple of BLK_G) and column
//     tiles of C positions (C Fp floats = 64 KB of LDS, Fp = padded frames);
//   * entries tile-major: for each tile, the segments of all row positions back
//     to back, each padded with zero entries to a multiple of 4 (16-byte loads);
//     an entry is (tile-local column, stored value), packed in 32 bits for
//     16-bit values;
//   * cnt[h m_pad + p]: padded length of row position p's segment in tile h;
//     seg_base[h (nG+1) + j]: first entry of granule j (BLK_G row positions) in
//     tile h, so that a row block's entries in a tile are one contiguous range;
//     nz_base[h (nG+1) + j]: row positions with entries in tile h before
//     granule j (a running count over tiles), so that a wavefront places its
//     granule's rows in the block's compacted row list without a block-wide
//     scan.
//
// A workgroup owns a row block and keeps its rows' F-frame sums in LDS while
// it walks the column tiles: for every tile the block touches it stages the
// tile's C columns x Fp frames of G in LDS, then shares the block's entries in
// that tile out over its slots of L lanes (L = Fp / P, P frames per lane).
// Only a quarter of a block's rows meet a given tile, so entries, not rows,
// are what is balanced: slot k takes the rows between the row boundaries
// nearest to the ends of its equal share of the tile's entry quads, walks
// their entries, and keeps the current row's sums in registers, read from the
// LDS row sums when the row starts and written back when it ends (no other
// slot touches the row in this tile).  Loads are requested a step ahead: the
// next tile's slice of G, counts and bases while the current tile is worked,
// and each slot's next batch of entry quads (within the run, or the first of
// its run in the next tile) while the current batch is.  The matrix streams
// once, no partial sums per (row, tile) reach memory, and the complete row
// sums feed the epilogue: the MLEM ratio and the FP64 log-likelihood for the
// forward projection, the image for the back projection.  When a direction
// has few row blocks (the back projection's voxel rows), each block's tile
// range is cut into `splits` equal parts for as many workgroups, which write
// partial sums that a reduction adds in split order.  The grid is persistent
// (one workgroup per CU looping over the work items), and the (rows per
// block, frames per lane) shape is autotuned per matrix and frame count.
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
constexpr int BLK_TILE_FLOATS = 16384;      // 64 KB of the gathered vector per tile
constexpr int BLK_G = 64;                   // row granule of the segment bases
constexpr int BLK_ENT_PAD = 64;             // zero entries after the stream (batches read past runs)
constexpr int BLK_MAX_RB = BLK_THREADS;     // rows per block (one row per thread)
constexpr int BLK_LDS_LIMIT = 160 * 1024 - 2048;   // LDS per workgroup (gfx950 has 160 KB)
// entry quads per batch of a slot; a lane holds two batches (the one it works
// on and the next, requested ahead), so fewer with 8 frames per lane, whose
// gathered rows take twice the registers, and with FP32 values, whose
// entries take 6 bytes (with 16-bit values and 4 frames per lane, 3 quads
// ran the 32-frame forward product 2% faster than 4)
template <typename VT, int P> constexpr int blk_u() {
  return sizeof(typename VT::T) == 4 ? (P == 4 ? 2 : 1) : (P == 4 ? 3 : 2);
}

// Blocked copy of one matrix for one padded frame count.
struct BlkCopy {
  int Fp = 0, lgFp = 0;        // padded frames (power of two, >= 8)
  int C = 0, lgC = 0, T = 0;   // tile columns, tiles
  int m = 0, m_pad = 0, nG = 0;
  int vt = EPET_VAL_F32;
  int* rowperm = nullptr;      // m_pad: position -> row, -1 past m
  int* colinv = nullptr;       // T C: position -> column, -1 past n
  int64_t* seg_base = nullptr; // T (nG + 1)
  int64_t* nz_base = nullptr;  // T (nG + 1)
  uint16_t* cnt = nullptr;     // T m_pad
  void* ent = nullptr;         // u32 packed (16-bit values) or u16 columns (FP32)
  float* eval = nullptr;       // FP32 values (FP32 storage)
  long long nent = 0;
  float* partial = nullptr;    // splits x m_pad x Fp
  long long partial_len = 0;
  float* gp = nullptr;         // the gathered vector in column-position order, T C x Fp
  int* work = nullptr;         // {next work item, finished workgroups}, zero between launches
  // launch shape (autotuned): frames per lane, rows per block, splits
  int P = 4, rb = 256, splits = 1;
  int tuned_F = 0;
  float ms = 0.0f;
  BlkCopy* next = nullptr;     // copies for other padded frame counts
  ~BlkCopy() {
    if (rowperm) (void)hipFree(rowperm);
    if (colinv) (void)hipFree(colinv);
    if (seg_base) (void)hipFree(seg_base);
    if (nz_base) (void)hipFree(nz_base);
    if (cnt) (void)hipFree(cnt);
    if (ent) (void)hipFree(ent);
    if (eval) (void)hipFree(eval);
    if (partial) (void)hipFree(partial);
    if (gp) (void)hipFree(gp);
    if (work) (void)hipFree(work);
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
  int L, S;                    // lanes per row slot, row slots per wavefront
  int lgNS;                    // log2 of the slots per workgroup
  int mode;                    // 0: forward MLEM ratio, 1: image, 2: partial sums
  const int* rowperm;
  const int64_t* seg_base;
  const int64_t* nz_base;
  const uint16_t* cnt;
  const void* ent;
  const float* eval;
  const float* Gp;             // gathered vector in column-position order, T C x Fp
  float* out;                  // forward: ratio in position order (m_pad x Fp);
                               // back projection: image (n x F, frame-minor)
  float* partial;
  int* work;                   // work-item counter and finished-workgroup count
  const float* rowscale;       // forward: per-row scale of the stored values
  int fold;                    // forward: the ratio carries the row scale
  const float* Y;              // forward: data (m x F)
  const int* itp;
  double* loglik;
};

// ------------------------------------------------------------------ entries
// Four consecutive entries (16-byte aligned), fetched raw and decoded when
// used: the columns as byte offsets of their frames in the 64 KB LDS tile (the
// 16-bit mask bounds every LDS address) and the values.
template <typename VT> struct BlkEnt;
template <> struct BlkEnt<ValBF16> {
  typedef uint4 Raw;
  static __device__ __forceinline__ Raw fetch(const BlkArgs& a, long long e) {
    return *reinterpret_cast<const uint4*>(static_cast<const uint32_t*>(a.ent) + e);
  }
  static __device__ __forceinline__ void decode(const Raw& pk, unsigned cm, unsigned c[4],
                                                float v[4]) {
    c[0] = pk.x & cm; c[1] = pk.y & cm; c[2] = pk.z & cm; c[3] = pk.w & cm;
    v[0] = __uint_as_float(pk.x & 0xFFFF0000u);
    v[1] = __uint_as_float(pk.y & 0xFFFF0000u);
    v[2] = __uint_as_float(pk.z & 0xFFFF0000u);
    v[3] = __uint_as_float(pk.w & 0xFFFF0000u);
  }
};
template <> struct BlkEnt<ValF16> {
  typedef uint4 Raw;
  static __device__ __forceinline__ Raw fetch(const BlkArgs& a, long long e) {
    return *reinterpret_cast<const uint4*>(static_cast<const uint32_t*>(a.ent) + e);
  }
  static __device__ __forceinline__ void decode(const Raw& pk, unsigned cm, unsigned c[4],
                                                float v[4]) {
    c[0] = pk.x & cm; c[1] = pk.y & cm; c[2] = pk.z & cm; c[3] = pk.w & cm;
    v[0] = __half2float(__ushort_as_half((uint16_t)(pk.x >> 16)));
    v[1] = __half2float(__ushort_as_half((uint16_t)(pk.y >> 16)));
    v[2] = __half2float(__ushort_as_half((uint16_t)(pk.z >> 16)));
    v[3] = __half2float(__ushort_as_half((uint16_t)(pk.w >> 16)));
    // one conversion per value, then full-rate FP32 FMAs (instead of the four
    // mixed-precision FMAs per value the compiler would otherwise fold them
    // into: 4 fewer instructions per quad, but the 32-frame forward product
    // ran 3.6% slower that way)
    asm volatile("" : "+v"(v[0]), "+v"(v[1]), "+v"(v[2]), "+v"(v[3]));
  }
};
typedef unsigned short blk_us4 __attribute__((ext_vector_type(4)));
struct BlkRawF32 {
  blk_us4 cc;
  float4 vv;
};
template <> struct BlkEnt<ValF32> {
  typedef BlkRawF32 Raw;
  static __device__ __forceinline__ Raw fetch(const BlkArgs& a, long long e) {
    Raw r;
    r.cc = *reinterpret_cast<const blk_us4*>(static_cast<const uint16_t*>(a.ent) + e);
    r.vv = *reinterpret_cast<const float4*>(a.eval + e);
    return r;
  }
  static __device__ __forceinline__ void decode(const Raw& r, unsigned cm, unsigned c[4],
                                                float v[4]) {
    c[0] = r.cc.x & cm; c[1] = r.cc.y & cm; c[2] = r.cc.z & cm; c[3] = r.cc.w & cm;
    v[0] = r.vv.x; v[1] = r.vv.y; v[2] = r.vv.z; v[3] = r.vv.w;
  }
};

// ------------------------------------------------------------------ main kernel
// LDS of a workgroup (one array of BLK_LDS_LIMIT bytes; the tile starts at
// LDS address 0, so an entry's byte offset is its LDS address):
//   tile      C Fp floats (64 KB) of Gp
//   acc       rb Fp floats: the block's row sums
//   lists     two sets, the current tile's and the next one's, each of
//             rs     rb + 4 int pairs: the block's rows with entries in the
//                    tile, compacted, in position order, as (LDS byte
//                    address of the row's sums, end of its segment) after a
//                    leading (0, 0), so that row k starts at rs[k].y,
//             slot   NS + 4 ints: each slot's first row (slot[NS] = rows),
//                    NS = the workgroup's slots;
//   tl_c0     BLK_TL int64, tl_h and tl_n BLK_TL ints: the non-empty tiles of
//             a chunk of BLK_TL tiles, the block's first entry and its entry
//             count in each;
//   wsum      2 x 16 + 2 ints; ll 64 doubles.
constexpr int BLK_TL = 512;                        // tiles per chunk of the tile list
__host__ __device__ inline int blk_list_ints(int rb, int ns) { return 2 * (rb + 4) + ns + 4; }
// slots of a workgroup: BLK_WAVES wavefronts of 64 / (Fp / P) slots
__host__ __device__ inline int blk_slots(int Fp, int P) { return BLK_WAVES * (WAVE / (Fp / P)); }
constexpr int BLK_LDS_TAIL = sizeof(int64_t) * BLK_TL + 2 * sizeof(int) * BLK_TL +
                             sizeof(int) * (2 * BLK_WAVES + 2) + sizeof(double) * 64;
__host__ __device__ inline size_t blk_lds_bytes(int C, int Fp, int rb, int ns) {
  return sizeof(float) * ((size_t)C * Fp + (size_t)rb * Fp) +
         sizeof(int) * 2 * (size_t)blk_list_ints(rb, ns) + BLK_LDS_TAIL;
}

// Workgroup-wide exclusive scan of (a, b) pairs in thread order: returns this
// thread's prefixes and the totals.  Two barriers.
__device__ __forceinline__ void blk_scan2(int xa, int xb, int* wsum, int* pa, int* pb, int* ta,
                                          int* tb) {
  const int lane = lane_id(), w = wave_id();
  int ya = xa, yb = xb;
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const int za = __shfl_up(ya, off, WAVE), zb = __shfl_up(yb, off, WAVE);
    if (lane >= off) { ya += za; yb += zb; }
  }
  if (lane == WAVE - 1) { wsum[w] = ya; wsum[BLK_WAVES + w] = yb; }
  __syncthreads();
  int ba = 0, bb = 0, sa = 0, sb = 0;
#pragma unroll
  for (int v = 0; v < BLK_WAVES; v++) {
    const int ca = wsum[v], cb = wsum[BLK_WAVES + v];
    ba += v < w ? ca : 0;
    bb += v < w ? cb : 0;
    sa += ca;
    sb += cb;
  }
  __syncthreads();
  *pa = ba + ya - xa;
  *pb = bb + yb - xb;
  *ta = sa;
  *tb = sb;
}

// Inclusive scan over a wavefront.
__device__ __forceinline__ int blk_wave_scan(int x) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const int z = __shfl_up(x, off, WAVE);
    if (lane >= off) x += z;
  }
  return x;
}

// One tile's row list (see the LDS layout above).
struct BlkList {
  int2* rs;   // rs[0] = (0, 0); rs[k + 1] = (LDS byte address of row k's sums, end of row k)
  int* slot;
};

// Places row t (cn entries from pc, compacted index pn, its sums at LDS byte
// address ra) of a tile of tc entries in list L, and marks the slots whose
// first row it is.  Slot k's equal share of the tile starts at entry es(k) =
// 4 floor(k (tc / 4) / NS); its run starts at the row boundary nearest to
// es(k) (the start of the row holding es(k) when es(k) lies in the row's first
// half, else its end), so that every row is one slot's, no two slots add into
// one row of the sums, and a run differs from the share by at most half a
// row at either end.  Row k starts where row k - 1 ends (rs[k].y).
__device__ __forceinline__ void blk_place(const BlkList& L, int ra, int cn, int pc, int pn, int tc,
                                          int lgNS) {
  if (cn <= 0) return;
  const int NS = 1 << lgNS;
  L.rs[pn + 1] = make_int2(ra, pc + cn);
  const unsigned nq = (unsigned)tc >> 2;
  // first slot whose share starts at or after pc
  unsigned k = (((unsigned)pc >> 2 << lgNS) + nq - 1) / nq;
  for (; k < (unsigned)NS; k++) {
    const int es = (int)(((nq * k) >> lgNS) << 2);
    if (es >= pc + cn) break;
    L.slot[k] = 2 * (es - pc) <= cn ? pn : pn + 1;
  }
  if (pc + cn == tc) L.slot[NS] = pn + 1;
}

template <typename VT, int P>
__global__ __launch_bounds__(BLK_THREADS) void k_blk(BlkArgs a) {
  // one static array, so that its address (0) is known when compiling and the
  // tile's byte offsets are LDS addresses as they are
  __shared__ float4 blk_lds[BLK_LDS_LIMIT / 16];
  typedef typename BlkEnt<VT>::Raw Raw;
  constexpr int U = blk_u<VT, P>();             // entry quads per batch
  float* xs = reinterpret_cast<float*>(blk_lds);
  float* acc = xs + ((size_t)a.C << a.lgFp);
  int* lists = reinterpret_cast<int*>(acc + ((size_t)a.rb << a.lgFp));
  const int li = blk_list_ints(a.rb, 1 << a.lgNS);
  // list set s (0 or 1)
  auto list_set = [&](int s) {
    BlkList L;
    L.rs = reinterpret_cast<int2*>(lists + s * li);
    L.slot = lists + s * li + 2 * (a.rb + 4);
    return L;
  };
  int64_t* tl_c0 = reinterpret_cast<int64_t*>(lists + 2 * li);
  int* tl_h = reinterpret_cast<int*>(tl_c0 + BLK_TL);
  int* tl_n = tl_h + BLK_TL;
  int* wsum = tl_n + BLK_TL;
  double* s_ll = reinterpret_cast<double*>(wsum + 2 * BLK_WAVES + 2);
  const int t = threadIdx.x;
  const int lane = lane_id();
  const int w = __builtin_amdgcn_readfirstlane(wave_id());
  const int slot = lane / a.L, q = lane % a.L;
  const int ks = w * a.S + slot;           // this lane's slot
  const unsigned cm = 0xFFFFu;
  const int rbfp = a.rb << a.lgFp;
  const int acc_addr = (a.C << a.lgFp) * (int)sizeof(float);   // LDS byte address of acc
  const int tile4 = (a.C << a.lgFp) >> 2;  // float4 per tile (4096)
  const unsigned lane_off = (unsigned)(q * P * 4);
  const char* xb = reinterpret_cast<const char*>(xs);
  const float4* g4 = reinterpret_cast<const float4*>(a.Gp);
  const unsigned long long lt_mask = (1ull << lane) - 1ull;
  if (t < 64) s_ll[t] = 0.0;
  if (t < 2) list_set(t).rs[0] = make_int2(0, 0);
  double ll = 0.0;                         // log-likelihood of frame t mod Fp
  int* s_item = wsum + 2 * BLK_WAVES;     // the work item this workgroup took
  const int nwork = a.nblocks * a.splits;
  // work items are taken in order from a counter, so a workgroup that finishes
  // early takes the next one
  for (;;) {
    __syncthreads();                       // the previous item's epilogue is done
    if (t == 0) *s_item = atomicAdd(a.work, 1);
    __syncthreads();
    const int work = *s_item;
    if (work >= nwork) break;
    const int blk = work / a.splits, sp = work % a.splits;
    const int p0 = blk * a.rb;
    const int rbe = min(a.rb, a.m_pad - p0);
    const int j0 = p0 / BLK_G, j1 = (p0 + rbe) / BLK_G;
    const bool wrows = w * WAVE < rbe;     // this wavefront holds rows (granule j0 + w)
    // this work item's tiles: an equal share of the tile range
    const int hlo = (int)((long long)a.T * sp / a.splits);
    const int hhi = (int)((long long)a.T * (sp + 1) / a.splits);
    for (int i = t; i < rbfp; i += BLK_THREADS) acc[i] = 0.0f;
    // a tile's row counts and granule bases for this block
    int cn = 0;
    long long sbw = 0, nzw = 0, nz0 = 0;
    auto request_meta = [&](int h) {
      cn = t < rbe ? (int)a.cnt[(long long)h * a.m_pad + p0 + t] : 0;
      if (wrows) {
        const long long o = (long long)h * (a.nG + 1);
        sbw = a.seg_base[o + j0 + w];
        nzw = a.nz_base[o + j0 + w];
        nz0 = a.nz_base[o + j0];
      }
    };
    // the row list of a tile (first entry c0x, tc entries) from the requested
    // counts and bases: a wavefront places its granule's rows
    auto build = [&](const BlkList& L, long long c0x, int tcx) {
      if (wrows) {
        const int pc = (int)(sbw - c0x) + blk_wave_scan(cn) - cn;
        const int pn = (int)(nzw - nz0) + __popcll(__ballot(cn > 0) & lt_mask);
        blk_place(L, acc_addr + (t << (a.lgFp + 2)), cn, pc, pn, tcx, a.lgNS);
      }
    };
    // this slot's rows and entries in a tile whose list is L
    auto run_of = [&](const BlkList& L, int* e0, int* e1) {
      *e0 = L.rs[L.slot[ks]].y;
      *e1 = L.rs[L.slot[ks + 1]].y;
    };
    for (int hb = hlo; hb < hhi; hb += BLK_TL) {
      // the block's non-empty tiles among the next BLK_TL, in order, with its
      // first entry and entry count in each
      const int hh = hb + t;
      long long c0h = 0;
      int nh = 0;
      if (t < BLK_TL && hh < hhi) {
        const int64_t* sb = a.seg_base + (long long)hh * (a.nG + 1);
        c0h = sb[j0];
        nh = (int)(sb[j1] - c0h);
      }
      int pk, pz, ntl, dz;
      __syncthreads();                     // the previous chunk's lists are no longer read
      blk_scan2(nh > 0 ? 1 : 0, 0, wsum, &pk, &pz, &ntl, &dz);
      if (nh > 0) {
        tl_h[pk] = hh;
        tl_c0[pk] = c0h;
        tl_n[pk] = nh;
      }
      __syncthreads();
      if (ntl == 0) continue;
      // tile 0's list, then the requests for its slice and tile 1's counts
      request_meta(tl_h[0]);
      build(list_set(0), tl_c0[0], tl_n[0]);
      long long gb = (long long)tl_h[0] * tile4 + t;
      float4 g0 = g4[gb], g1 = g4[gb + BLK_THREADS], g2 = g4[gb + 2 * BLK_THREADS],
             g3 = g4[gb + 3 * BLK_THREADS];
      request_meta(tl_h[min(1, ntl - 1)]);
      __syncthreads();                     // tile 0's list is complete
      // bA holds the first batch of this slot's run in the current tile
      Raw bA[U];
      {
        int e, e_end;
        run_of(list_set(0), &e, &e_end);
#pragma unroll
        for (int u = 0; u < U; u++)
          bA[u] = BlkEnt<VT>::fetch(a, tl_c0[0] + e + 4 * u);
      }
      for (int it = 0; it < ntl; it++) {
        const long long c0 = tl_c0[it];
        const bool has_next = it + 1 < ntl;
        const BlkList Lc = list_set(it & 1);
        const BlkList Ln = list_set((it + 1) & 1);
        __syncthreads();                   // the previous tile's readers are done
        float4* x4 = reinterpret_cast<float4*>(xs);
        x4[t] = g0;
        x4[t + BLK_THREADS] = g1;
        x4[t + 2 * BLK_THREADS] = g2;
        x4[t + 3 * BLK_THREADS] = g3;
        // the next tile's list, then the requests for its slice and the counts
        // of the one after
        if (has_next) build(Ln, tl_c0[it + 1], tl_n[it + 1]);
        {
          const int hn = tl_h[min(it + 1, ntl - 1)];
          gb = (long long)hn * tile4 + t;
          g0 = g4[gb];
          g1 = g4[gb + BLK_THREADS];
          g2 = g4[gb + 2 * BLK_THREADS];
          g3 = g4[gb + 3 * BLK_THREADS];
          request_meta(tl_h[min(it + 2, ntl - 1)]);
        }
        __syncthreads();
        int e, e_end;
        run_of(Lc, &e, &e_end);
        // where this slot's batch requests go once its run here is done: the
        // first batch of its run in the next tile (after the last tile, any
        // in-bounds address: a batch may read up to 4 U - 1 entries past a
        // run, which the padding after the entries keeps inside the array)
        long long nxt = c0 + e_end;
        if (has_next) {
          int en, en_end;
          run_of(Ln, &en, &en_end);
          nxt = tl_c0[it + 1] + en;
        }
        if (e >= e_end) {
          // nothing to do here: request the next tile's first batch now
#pragma unroll
          for (int u = 0; u < U; u++) bA[u] = BlkEnt<VT>::fetch(a, nxt + 4 * u);
        }
        // the current row (LDS byte address of its sums, end), the next one
        // read ahead, and the current row's sums, read from the LDS row sums
        // when the row starts, accumulated in registers and written back when
        // it ends (one slot owns a row within a tile)
        const int k0 = Lc.slot[ks];
        const int2 c0r = Lc.rs[k0 + 1];
        unsigned ra = (unsigned)c0r.x + lane_off;   // this lane's part of the row's sums
        int rend = c0r.y;
        int2 nx = Lc.rs[k0 + 2];
        const int2* kp = Lc.rs + k0 + 3;
        float a4[P];
        auto load_row = [&]() {
#pragma unroll
          for (int i = 0; i < P / 4; i++) {
            const float4 o = *reinterpret_cast<const float4*>(xb + ra + 16 * i);
            a4[4 * i + 0] = o.x;
            a4[4 * i + 1] = o.y;
            a4[4 * i + 2] = o.z;
            a4[4 * i + 3] = o.w;
          }
        };
        if (e < e_end) load_row();
        // works the quads of batch b that lie in the run
        auto work = [&](const Raw (&b)[U]) {
#pragma unroll
          for (int u = 0; u < U; u++) {
            if (e >= e_end) break;             // quads past the run are not worked
            unsigned c[4];
            float v[4];
            BlkEnt<VT>::decode(b[u], cm, c, v);
            // the quad's four gathered rows requested together, then accumulated
            float4 xv[4][P / 4];
#pragma unroll
            for (int t4 = 0; t4 < 4; t4++) {
              const float4* xp = reinterpret_cast<const float4*>(xb + (c[t4] | lane_off));
#pragma unroll
              for (int i = 0; i < P / 4; i++) xv[t4][i] = xp[i];
            }
#pragma unroll
            for (int t4 = 0; t4 < 4; t4++)
#pragma unroll
              for (int i = 0; i < P / 4; i++) {
                a4[4 * i + 0] += v[t4] * xv[t4][i].x;
                a4[4 * i + 1] += v[t4] * xv[t4][i].y;
                a4[4 * i + 2] += v[t4] * xv[t4][i].z;
                a4[4 * i + 3] += v[t4] * xv[t4][i].w;
              }
            e += 4;
            if (e == rend) {
              // the row is complete in this tile
              float4* dst = reinterpret_cast<float4*>(const_cast<char*>(xb) + ra);
#pragma unroll
              for (int i = 0; i < P / 4; i++)
                dst[i] = make_float4(a4[4 * i + 0], a4[4 * i + 1], a4[4 * i + 2], a4[4 * i + 3]);
              if (e < e_end) {
                ra = (unsigned)nx.x + lane_off;
                rend = nx.y;
                nx = *kp++;
                load_row();
              }
            }
          }
        };
        // the batch after the current one (starting at e): the run's next, or
        // the next tile's first once the run ends in the current one
        auto next_at = [&]() { return e + 4 * U < e_end ? c0 + e + 4 * U : nxt; };
        // two batch buffers in turn, so that the batch requested ahead lands
        // in the buffer not being worked (no register copies per batch); the
        // loop leaves only at the top, with the next tile's first batch in bA
        Raw bB[U];
        for (;;) {
          if (!__any(e < e_end)) break;
          {
            const long long at = next_at();
#pragma unroll
            for (int u = 0; u < U; u++) bB[u] = BlkEnt<VT>::fetch(a, at + 4 * u);
          }
          if (e < e_end) work(bA);
          {
            const long long at = next_at();
#pragma unroll
            for (int u = 0; u < U; u++) bA[u] = BlkEnt<VT>::fetch(a, at + 4 * u);
          }
          if (e < e_end) work(bB);
        }
      }
    }
    __syncthreads();
    // ---- epilogue: the block's rows are complete (or this split's share is)
    if (a.mode == 2) {
      float* dst = a.partial + (((long long)sp * a.m_pad + p0) << a.lgFp);
      for (int i = t; i < (rbe << a.lgFp) / 4; i += BLK_THREADS)
        reinterpret_cast<float4*>(dst)[i] = reinterpret_cast<const float4*>(acc)[i];
      continue;
    }
    for (int i = t; i < (rbe << a.lgFp); i += BLK_THREADS) {
      const int r = i >> a.lgFp, f = i & (a.Fp - 1);
      if (f >= a.F) continue;
      const int orig = a.rowperm[p0 + r];
      if (orig < 0) continue;
      if (a.mode == 1) {
        a.out[(long long)orig * a.F + f] = acc[i];
      } else {
        const float sc = a.rowscale ? a.rowscale[orig] : 1.0f;
        const float yh = acc[i] * sc;
        const float yi = a.Y[(long long)orig * a.F + f];
        const float ev = yh > 0.0f ? yi / yh : 0.0f;
        a.out[((long long)(p0 + r) << a.lgFp) + f] = a.fold ? ev * sc : ev;
        if (yh > 0.0f) ll += (double)yi * log((double)yh) - (double)yh;
      }
    }
  }
  if (a.mode == 0 && a.loglik) {
    // thread t only ever saw frame t mod Fp
    __syncthreads();
    atomicAdd(&s_ll[t & (a.Fp - 1)], ll);
    __syncthreads();
    if (t < a.F) unsafeAtomicAdd(&a.loglik[(long long)(*a.itp) * a.F + t], s_ll[t]);
  }
  // the last workgroup to finish zeroes the counters for the next launch
  if (t == 0) {
    __threadfence();
    if (atomicAdd(a.work + 1, 1) == (int)gridDim.x - 1) {
      a.work[0] = 0;
      a.work[1] = 0;
    }
  }
}

// Gp[pos Fp + f] = G[colinv[pos] F + f] (0 for padding positions and frames
// past F): the gathered vector in the copy's column order, so that a tile is
// one contiguous 64 KB read.
__global__ __launch_bounds__(256) void k_blk_gather(float* __restrict__ Gp,
                                                    const float* __restrict__ G,
                                                    const int* __restrict__ colinv, long long npos,
                                                    int F, int lgFp) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (npos << lgFp)) return;
  const long long pos = i >> lgFp;
  const int f = (int)(i & ((1 << lgFp) - 1));
  const int c = colinv[pos];
  Gp[i] = (c >= 0 && f < F) ? G[(long long)c * F + f] : 0.0f;
}

// Split reductions, in split order.  Forward: the MLEM ratio (in row-position
// order) and the FP64 log-likelihood (one atomic per workgroup and frame); back
// projection: the image.
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
      a.out[(p << a.lgFp) + f] = a.fold ? ev * sc : ev;
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
__global__ __launch_bounds__(256) void k_blk_fill(int m_pad, int T, int lgC, int lgFp,
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
          // the entry's column as the byte offset of its frames in the LDS tile
          // (C Fp floats = 64 KB, so it fits 16 bits)
          const unsigned local = ((unsigned)pos & cm) << (lgFp + 2);
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

// zero entries [at, at + BLK_ENT_PAD) after the stream
__global__ void k_blk_zero_tail(void* ent, float* eval, long long at, int eb) {
  const int i = threadIdx.x;
  if (i < BLK_ENT_PAD) {
    if (eb == 4) static_cast<uint32_t*>(ent)[at + i] = 0u;
    else static_cast<uint16_t*>(ent)[at + i] = 0;
    if (eval) eval[at + i] = 0.0f;
  }
}

__global__ __launch_bounds__(256) void k_blk_narrow(const int* __restrict__ c32, long long n,
                                                    uint16_t* __restrict__ c16) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c16[i] = (uint16_t)c32[i];
}

// sb[h (nG+1) + j] = off[h stride + j step] (j = nG: the tile's end, which is
// the next tile's start)
__global__ __launch_bounds__(256) void k_blk_bases(int T, int nG, long long stride, int step,
                                                   const int64_t* __restrict__ off,
                                                   int64_t* __restrict__ sb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)T * (nG + 1)) return;
  const long long h = i / (nG + 1), j = i % (nG + 1);
  sb[i] = off[h * stride + j * step];
}

// gnz[h nG + j]: row positions of granule j with entries in tile h (one
// wavefront per granule)
__global__ __launch_bounds__(256) void k_blk_granule_nz(long long ngran, const int* __restrict__ c32,
                                                        int* __restrict__ gnz) {
  const long long g = (long long)blockIdx.x * (256 / WAVE) + wave_id();
  if (g >= ngran) return;
  const unsigned long long b = __ballot(c32[g * BLK_G + lane_id()] > 0);
  if (lane_id() == 0) gnz[g] = __popcll(b);
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
  const long long ngp = ((long long)B->T * B->C) << B->lgFp;
  bool ok = hipMalloc((void**)&B->gp, sizeof(float) * (size_t)ngp) == hipSuccess &&
            hipMalloc((void**)&B->rowperm, sizeof(int) * (size_t)B->m_pad) == hipSuccess &&
            hipMalloc((void**)&B->work, sizeof(int) * 2) == hipSuccess &&
            hipMalloc((void**)&B->colinv, sizeof(int) * ci.size()) == hipSuccess &&
            hipMalloc((void**)&cpos_d, sizeof(int) * cpos.size()) == hipSuccess &&
            hipMalloc((void**)&B->seg_base, sizeof(int64_t) * (size_t)B->T * (B->nG + 1)) ==
                hipSuccess &&
            hipMalloc((void**)&B->nz_base, sizeof(int64_t) * (size_t)B->T * (B->nG + 1)) ==
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
  int rc = EPET_OK;
  if (hipMemcpyAsync(B->rowperm, rp.data(), sizeof(int) * rp.size(), hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      hipMemcpyAsync(B->colinv, ci.data(), sizeof(int) * ci.size(), hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      hipMemcpyAsync(cpos_d, cpos.data(), sizeof(int) * cpos.size(), hipMemcpyHostToDevice, st) !=
          hipSuccess) {
    rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "blocked copy: orders");
  } else {
    k_fill_i32_v<<<grid_for(nseg, 256, 1ll << 31), 256, 0, st>>>(cnt32, nseg, 0);
    k_fill_i32_v<<<1, 64, 0, st>>>(B->work, 2, 0);
    k_fill_i32_v<<<grid_for(ngp, 256, 1ll << 31), 256, 0, st>>>(reinterpret_cast<int*>(B->gp), ngp,
                                                                0);
  }
  const size_t lds = sizeof(int) * (256 / WAVE) * (size_t)B->T;
  if (rc == EPET_OK && lds > 160 * 1024) {
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
        B->T, B->nG, B->m_pad, BLK_G, off, B->seg_base);
    // row positions with entries per (tile, granule), scanned over tiles and
    // granules into nz_base
    const long long ngran = (long long)B->T * B->nG;
    int* gcount = nullptr;
    int64_t* nzoff = nullptr;
    if (hipMalloc((void**)&gcount, sizeof(int) * (size_t)std::max(1ll, ngran)) != hipSuccess ||
        hipMalloc((void**)&nzoff, sizeof(int64_t) * (size_t)(ngran + 1)) != hipSuccess) {
      (void)hipGetLastError();
      if (gcount) (void)hipFree(gcount);
      cleanup();
      delete B;
      return EPET_OK;
    }
    k_blk_granule_nz<<<grid_for(ngran, 256 / WAVE, 1ll << 31), 256, 0, st>>>(ngran, cnt32, gcount);
    long long nzt = 0;
    rc = exclusive_scan_i32_to_i64(ctx, gcount, nzoff, ngran, &nzt);
    if (rc == EPET_OK) {
      k_set_i64_at<<<1, 64, 0, st>>>(nzoff + ngran, nzt);
      k_blk_bases<<<grid_for((long long)B->T * (B->nG + 1), 256, 1ll << 31), 256, 0, st>>>(
          B->T, B->nG, B->nG, 1, nzoff, B->nz_base);
      if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "blocked copy row bases");
    }
    (void)hipFree(gcount);
    (void)hipFree(nzoff);
    (void)hipFree(cnt32);
    cnt32 = nullptr;
    if (rc != EPET_OK) {
      cleanup();
      delete B;
      return rc;
    }
    // entries, then BLK_ENT_PAD zero entries that batches reading past a run reach
    const size_t eb = M->value_type == EPET_VAL_F32 ? 2 : 4;
    if (hipMalloc(&B->ent, eb * (size_t)(tot + BLK_ENT_PAD)) != hipSuccess ||
        (M->value_type == EPET_VAL_F32 &&
         hipMalloc((void**)&B->eval, sizeof(float) * (size_t)(tot + BLK_ENT_PAD)) != hipSuccess)) {
      (void)hipGetLastError();
      cleanup();
      delete B;
      return EPET_OK;
    }
    k_blk_zero_tail<<<1, BLK_ENT_PAD, 0, st>>>(B->ent, B->eval, tot, (int)eb);
    if (M->value_type == EPET_VAL_F32)
      k_blk_fill<float><<<g, 256, lds, st>>>(B->m_pad, B->T, B->lgC, B->lgFp, B->rowperm, M->rowptr, M->col,
                                             (const float*)M->val, cpos_d, off, B->ent, B->eval);
    else
      k_blk_fill<uint16_t><<<g, 256, lds, st>>>(B->m_pad, B->T, B->lgC, B->lgFp, B->rowperm, M->rowptr,
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

static BlkKern blk_kernel_vt(int vt, int P) {
#define EPET_BK(VTT)                          \
  switch (P) {                                \
    case 4: return k_blk<VTT, 4>;             \
    case 8: return k_blk<VTT, 8>;             \
    default: return nullptr;                  \
  }
  switch (vt) {
    case EPET_VAL_F32: EPET_BK(ValF32)
    case EPET_VAL_BF16: EPET_BK(ValBF16)
    default: EPET_BK(ValF16)
  }
#undef EPET_BK
}

// A (frames per lane, rows per block) shape fits when the lanes per slot
// divide a wavefront and the workgroup's LDS holds the tile and the row sums.
static bool blk_shape_valid(const BlkCopy* B, int P, int rb) {
  if (P > B->Fp || B->Fp / P > WAVE || rb % BLK_G != 0 || rb > BLK_MAX_RB) return false;
  return blk_lds_bytes(B->C, B->Fp, rb, blk_slots(B->Fp, P)) <= (size_t)BLK_LDS_LIMIT;
}

// work-item splits of a direction: eight work items per CU at least, so that
// the items left when the first workgroups run out of work are small
static int blk_splits(const BlkCopy* B, int rb, int ncu) {
  const int nblocks = (B->m_pad + rb - 1) / rb;
  const int want = 8 * ncu;
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

// One product with an explicit shape over B's position-ordered gathered
// vector B->gp (mode 0: forward MLEM step writing the ratio in row-position
// order to `out`; mode 1: image).
static int blk_launch(epet_ctx* ctx, const BlkCopy* B, int P, int rb, int splits, int F,
                      float* out, int mode, const float* rowscale, int fold, const float* Y,
                      const int* itp, double* loglik) {
  BlkKern kern = blk_kernel_vt(B->vt, P);
  EPET_CHECK(kern && blk_shape_valid(B, P, rb), EPET_ERR_STATE,
             "blocked products: no kernel for this shape");
  EPET_CHECK(splits <= 1 || (B->partial && B->partial_len >= (long long)splits * B->m_pad * B->Fp),
             EPET_ERR_STATE, "blocked products: partial buffer missing");
  BlkArgs a{};
  a.F = F;
  a.Fp = B->Fp;
  a.lgFp = B->lgFp;
  a.C = B->C;
  a.T = B->T;
  a.m_pad = B->m_pad;
  a.nG = B->nG;
  a.rb = rb;
  a.nblocks = (B->m_pad + rb - 1) / rb;
  a.splits = splits;
  a.L = B->Fp / P;
  a.S = WAVE / a.L;
  a.lgNS = ilog2(BLK_WAVES * a.S);
  a.mode = splits > 1 ? 2 : mode;
  a.rowperm = B->rowperm;
  a.seg_base = B->seg_base;
  a.nz_base = B->nz_base;
  a.cnt = B->cnt;
  a.ent = B->ent;
  a.eval = B->eval;
  a.Gp = B->gp;
  a.out = out;
  a.partial = B->partial;
  a.work = B->work;
  a.rowscale = rowscale;
  a.fold = fold;
  a.Y = Y;
  a.itp = itp;
  a.loglik = loglik;
  const long long nwork = (long long)a.nblocks * a.splits;
  const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(nwork, ctx->num_cus));
  kern<<<grid, BLK_THREADS, 0, ctx->stream>>>(a);
  if (splits > 1) {
    a.mode = mode;
    k_blk_reduce<<<grid_for((long long)B->m_pad * F, 256, 16ll * ctx->num_cus), 256, 0,
                   ctx->stream>>>(a);
  }
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

// The forward copy's gathered vector (the image) in its column order.  The
// back projection's gathered vector (the ratio) is written in that order by the
// forward product itself: the forward copy's row order is the transposed
// copy's column order.
static int blk_gather_x(epet_ctx* ctx, const BlkCopy* f, int F, const float* x) {
  const long long npos = (long long)f->T * f->C;
  k_blk_gather<<<grid_for(npos << f->lgFp, 256, 1ll << 31), 256, 0, ctx->stream>>>(
      f->gp, x, f->colinv, npos, F, f->lgFp);
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
  EPET_TRY(blk_gather_x(ctx, f, a.F, a.x));
  EPET_TRY(blk_launch(ctx, f, f->P, f->rb, f->splits, a.F, b->gp, 0, A->rowscale,
                      A->rowscale ? 1 : 0, a.y, a.itp, a.loglik));
  EPET_TRY(blk_launch(ctx, b, b->P, b->rb, b->splits, a.F, a.c, 1, nullptr, 0, nullptr,
                      nullptr, nullptr));
  return EPET_OK;
}

// Autotuner: every valid (frames per lane P, rows per block rb) shape of each
// direction is timed with HIP events on the context's stream (reps launches
// after one warm-up) and the fastest kept.
int blk_tune(epet_ctx* ctx, epet_matrix* A, const MfArgs& a, int reps) {
  const int Fp = padded_frames(a.F);
  BlkCopy* dirs[2] = {blk_find(A, Fp), blk_find(A->tr, Fp)};
  EPET_CHECK(dirs[0] && dirs[1], EPET_ERR_STATE, "blocked copies missing");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return fail(EPET_ERR_HIP, __FILE__, __LINE__, "tuning events");
  }
  // frames per lane: 4 (a slot of Fp/4 lanes reads whole voxel rows, no bank
  // conflicts) or 8; 16 and 32 lost to bank conflicts at every block size
  // (53 and 86 ms against 35 ms per 32-frame forward product at 52^3)
  static const int Ps[2] = {4, 8};
  static const int RBs[4] = {128, 256, 512, 1024};
  int rc = blk_gather_x(ctx, dirs[0], a.F, a.x);
  for (int d = 0; d < 2 && rc == EPET_OK; d++) {
    BlkCopy* B = dirs[d];
    float best = 1e30f;
    int bP = 0, bR = 0, bS = 1;
    for (int pi = 0; pi < 2 && rc == EPET_OK; pi++)
      for (int ri = 0; ri < 4 && rc == EPET_OK; ri++) {
        const int P = Ps[pi], rb = RBs[ri];
        if (!blk_shape_valid(B, P, rb)) continue;
        const int splits = blk_splits(B, rb, ctx->num_cus);
        if (blk_partial_ensure(B, splits) != EPET_OK) { (void)hipGetLastError(); continue; }
        auto run = [&]() {
          return d == 0 ? blk_launch(ctx, B, P, rb, splits, a.F, dirs[1]->gp, 0, A->rowscale,
                                     A->rowscale ? 1 : 0, a.y, a.itp, nullptr)
                        : blk_launch(ctx, B, P, rb, splits, a.F, a.c, 1, nullptr, 0, nullptr,
                                     nullptr, nullptr);
        };
        rc = run();
        if (rc != EPET_OK) break;
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) {
          rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "event");
          break;
        }
        for (int r = 0; r < reps && rc == EPET_OK; r++) rc = run();
        if (rc != EPET_OK) break;
        float ms = 0.0f;
        if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "tuning events");
          break;
        }
        ms /= reps;
        if (ms < best) { best = ms; bP = P; bR = rb; bS = splits; }
      }
    if (rc == EPET_OK) {
      if (bP == 0) rc = fail(EPET_ERR_STATE, __FILE__, __LINE__, "blocked products: no valid shape");
      else {
        B->P = bP;
        B->rb = bR;
        B->splits = bS;
        B->ms = best;
        B->tuned_F = a.F;
        rc = blk_partial_ensure(B, bS);
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// Launch shape of the blocked products for F frames: {1, P_fwd, lanes per
// slot fwd, rows per block fwd, splits fwd, P_bwd, lanes per slot bwd, rows
// per block bwd, splits bwd} and the tuned times, or {0} when the blocked path
// is not in use.
int blk_info(const epet_matrix* A, int F, int info[9], float ms[2]) {
  for (int i = 0; i < 9; i++) info[i] = 0;
  ms[0] = ms[1] = 0.0f;
  if (!blk_ready(A, F)) return EPET_OK;
  const int Fp = padded_frames(F);
  const BlkCopy* f = blk_find(A, Fp);
  const BlkCopy* b = blk_find(A->tr, Fp);
  info[0] = 1;
  info[1] = f->P; info[2] = Fp / f->P; info[3] = f->rb; info[4] = f->splits;
  info[5] = b->P; info[6] = Fp / b->P; info[7] = b->rb; info[8] = b->splits;
  ms[0] = f->ms;
  ms[1] = b->ms;
  return EPET_OK;
}

}  // namespace epet
