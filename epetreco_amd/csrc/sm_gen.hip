// sm_gen.hip -- Monte-Carlo rays, workqueue and system-matrix generation.
//
// Compiled with -ffp-contract=off and correctly rounded division / square root:
// every value is the reference's single-precision arithmetic in the reference's
// order (see epet_geom.hpp), so device output is bit-identical with the CPU
// restatement given the same rays.
#include <algorithm>
#include <climits>
#include <vector>

#include "epet_block.hpp"
#include "epet_geom.hpp"

#pragma clang fp contract(off)

namespace epet {

constexpr int GEN_BLOCK = 256;
constexpr int MAX_SLABS = 512;        // x planes of the traversal path
constexpr int BITMAP_WORDS = 4096;    // 131072 voxel bits per row in LDS
constexpr float MARK_EPS = 2e-3f;     // traversal margin, in voxel widths
constexpr float REJECT_MARGIN2 = 1.0404f;  // (1.02)^2 on the voxel half diagonal

// ------------------------------------------------------------------ rays
__global__ __launch_bounds__(256) void k_rays(Geo g, const int* __restrict__ cnl, int n,
                                              float* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * g.nrays) return;
  const int i = (int)(t / g.nrays), r = (int)(t % g.nrays);
  const int c = cnl[i];
  const ChannelIds ci = channel_ids(g, c);
  float ray[6];
  make_ray(g, ci, c, r, ray);
#pragma unroll
  for (int k = 0; k < 6; k++) out[t * 6 + k] = ray[k];
}

// ------------------------------------------------------------------ workqueue
// One workgroup per list entry: candidate count from voxel `v0` of the first
// row (the cursor), from 0 for every later row.
__global__ __launch_bounds__(GEN_BLOCK) void k_wq_count(Geo g, const int* __restrict__ list,
                                                       int row0, int v0,
                                                       int* __restrict__ counts) {
  __shared__ int s_wave[GEN_BLOCK / WAVE];
  const int row = row0 + blockIdx.x;
  const ChannelIds ci = channel_ids(g, list[row]);
  float p0[3], p1[3];
  pixel_centres(g, ci, p0, p1);
  const int vstart = (blockIdx.x == 0) ? v0 : 0;
  int cnt = 0;
  for (int v = vstart + threadIdx.x; v < g.V; v += GEN_BLOCK) cnt += is_candidate(g, p0, p1, v);
  int tot;
  (void)block_excl_sum<GEN_BLOCK>(cnt, s_wave, &tot);
  if (threadIdx.x == 0) counts[blockIdx.x] = tot;
}

// Ordered emission at offsets[row] (exclusive scan of the counts), writing only
// the first `cap` pairs overall.
__global__ __launch_bounds__(GEN_BLOCK) void k_wq_emit(Geo g, const int* __restrict__ list,
                                                      int row0, int v0,
                                                      const int64_t* __restrict__ offsets,
                                                      long long cap, int* __restrict__ out_cnl,
                                                      int* __restrict__ out_vxl) {
  __shared__ int s_wave[GEN_BLOCK / WAVE];
  const int row = row0 + blockIdx.x;
  const int cnl = list[row];
  long long pos = offsets[blockIdx.x];
  if (pos >= cap) return;   // uniform across the workgroup
  const ChannelIds ci = channel_ids(g, cnl);
  float p0[3], p1[3];
  pixel_centres(g, ci, p0, p1);
  const int vstart = (blockIdx.x == 0) ? v0 : 0;
  for (int base = vstart; base < g.V; base += GEN_BLOCK) {
    const int v = base + threadIdx.x;
    const bool hit = v < g.V && is_candidate(g, p0, p1, v);
    int tot;
    const int r = block_rank<GEN_BLOCK>(hit, s_wave, &tot);
    if (hit && pos + r < cap) {
      out_cnl[pos + r] = cnl;
      out_vxl[pos + r] = v;
    }
    pos += tot;
    if (pos >= cap) break;  // uniform
  }
}

// ------------------------------------------------------------------ values
// getSystemMatrixFromWorkqueue: one thread per entry, rays replayed per entry.
__global__ __launch_bounds__(256) void k_sm_from_wq(Geo g, const int* __restrict__ cnl,
                                                   const int* __restrict__ vxl,
                                                   float* __restrict__ val, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cnl[i];
  const ChannelIds ci = channel_ids(g, c);
  float box[6];
  voxel_box(g, vxl[i], box);
  float a = 0.0f;
  for (int r = 0; r < g.nrays; r++) {
    float ray[6];
    make_ray(g, ci, c, r, ray);
    const float den[3] = {ray[3] - ray[0], ray[4] - ray[1], ray[5] - ray[2]};
    const float len = absolute3(den[0], den[1], den[2]);
    a += slab_length_masked(box, ray, den, len);
  }
  val[i] = a / (float)g.nrays;
}

// ------------------------------------------------------------------ fused rows
struct RowRays {
  float p[3][EPET_MAX_RAYS];    // start point
  float den[3][EPET_MAX_RAYS];  // end - start
  float u[3][EPET_MAX_RAYS];    // unit direction (conservative test only)
  float len[EPET_MAX_RAYS];     // |end - start|, reference expression
};

// a_ij = sum_r intersectionLength(voxel v, ray r) / nrays, with rays whose line
// passes farther than 1.02 voxel half-diagonals from the voxel centre skipped:
// their exact length is 0, so skipping them leaves the float sum unchanged.
__device__ __forceinline__ float row_value(const Geo& g, const RowRays& R, int nrays,
                                           bool degenerate, int v) {
  float box[6];
  voxel_box(g, v, box);
  const float c0 = 0.5f * (box[0] + box[3]);
  const float c1 = 0.5f * (box[1] + box[4]);
  const float c2 = 0.5f * (box[2] + box[5]);
  const float thr = g.halfDiag2 * REJECT_MARGIN2;
  float a = 0.0f;
  for (int r = 0; r < nrays; r++) {
    const float p0 = R.p[0][r], p1 = R.p[1][r], p2 = R.p[2][r];
    if (!degenerate) {
      const float w0 = c0 - p0, w1 = c1 - p1, w2 = c2 - p2;
      const float u0 = R.u[0][r], u1 = R.u[1][r], u2 = R.u[2][r];
      const float x0 = w1 * u2 - w2 * u1;
      const float x1 = w2 * u0 - w0 * u2;
      const float x2 = w0 * u1 - w1 * u0;
      if ((x0 * x0 + x1 * x1) + x2 * x2 > thr) continue;
      a += slab_length(box, p0, p1, p2, R.den[0][r], R.den[1][r], R.den[2][r], R.len[r]);
    } else {
      const float p[3] = {p0, p1, p2};
      const float den[3] = {R.den[0][r], R.den[1][r], R.den[2][r]};
      a += slab_length_masked(box, p, den, R.len[r]);
    }
  }
  return a / (float)nrays;
}

// Voxel rectangle [iy0,iy1] x [iz0,iz1] that ray r may cross inside x-plane i.
__device__ __forceinline__ bool slab_rect(const Geo& g, const RowRays& R, int r, int i,
                                          int& iy0, int& iy1, int& iz0, int& iz1) {
  const float xlo = g.go[0] + (float)i * g.gd[0];
  const float xhi = g.go[0] + (float)(i + 1) * g.gd[0];
  const float px = R.p[0][r], dx = R.den[0][r];
  float ta = (xlo - px) / dx, tb = (xhi - px) / dx;
  if (tb < ta) { const float t = ta; ta = tb; tb = t; }
  const float ya = R.p[1][r] + ta * R.den[1][r], yb = R.p[1][r] + tb * R.den[1][r];
  const float za = R.p[2][r] + ta * R.den[2][r], zb = R.p[2][r] + tb * R.den[2][r];
  const float fy0 = (fminf(ya, yb) - g.go[1]) / g.gd[1] - MARK_EPS;
  const float fy1 = (fmaxf(ya, yb) - g.go[1]) / g.gd[1] + MARK_EPS;
  const float fz0 = (fminf(za, zb) - g.go[2]) / g.gd[2] - MARK_EPS;
  const float fz1 = (fmaxf(za, zb) - g.go[2]) / g.gd[2] + MARK_EPS;
  if (fy1 < 0.0f || fz1 < 0.0f || fy0 >= (float)g.gn[1] || fz0 >= (float)g.gn[2]) return false;
  iy0 = max(0, (int)floorf(fy0));
  iy1 = min(g.gn[1] - 1, (int)floorf(fy1));
  iz0 = max(0, (int)floorf(fz0));
  iz1 = min(g.gn[2] - 1, (int)floorf(fz1));
  return iy0 <= iy1 && iz0 <= iz1;
}

struct RowShared {
  RowRays rays;
  int wy0[MAX_SLABS], wy1[MAX_SLABS], wz0[MAX_SLABS], wz1[MAX_SLABS];
  int off[MAX_SLABS + 1];
  uint32_t bits[BITMAP_WORDS];
  int wave[GEN_BLOCK / WAVE];
  int flag;
};

// One workgroup per listed channel.  Output: the row's (voxel, value) pairs in
// ascending voxel order at scratch + blockIdx.x * rowcap, and its count in
// row_cnt (negative: -needed when rowcap is too small).
//
// Nonzero mode (keep_zeros = 0): traversal marks every voxel within MARK_EPS of
// any ray (x-plane by x-plane, a bitmap over per-plane y/z windows), then the
// exact value and candidate predicate are evaluated for the marked voxels only.
// A voxel no ray comes near has exact value 0, so the result equals the
// reference's candidate set filtered by value != 0.  Rows whose rays are
// parallel to an axis (the reference then ignores that axis), grids wider than
// MAX_SLABS or windows larger than the bitmap use the full voxel scan.
// Candidate mode (keep_zeros = 1): full scan, every candidate kept.
__global__ __launch_bounds__(GEN_BLOCK) void k_sm_rows(Geo g, const int* __restrict__ list,
                                                      int row0, int rowcap, int keep_zeros,
                                                      int* __restrict__ scol,
                                                      float* __restrict__ sval,
                                                      int* __restrict__ row_cnt) {
  __shared__ RowShared S;
  const int tid = threadIdx.x;
  const int row = row0 + blockIdx.x;
  const int cnl = list ? list[row] : row;
  const int nrays = g.nrays;
  const ChannelIds ci = channel_ids(g, cnl);
  int* ocol = scol + (size_t)blockIdx.x * rowcap;
  float* oval = sval + (size_t)blockIdx.x * rowcap;

  if (tid == 0) S.flag = 0;
  __syncthreads();
  for (int r = tid; r < nrays; r += GEN_BLOCK) {
    float ray[6];
    make_ray(g, ci, cnl, r, ray);
    const float d0 = ray[3] - ray[0], d1 = ray[4] - ray[1], d2 = ray[5] - ray[2];
    const float len = absolute3(d0, d1, d2);
    S.rays.p[0][r] = ray[0]; S.rays.p[1][r] = ray[1]; S.rays.p[2][r] = ray[2];
    S.rays.den[0][r] = d0; S.rays.den[1][r] = d1; S.rays.den[2][r] = d2;
    S.rays.len[r] = len;
    const float il = len > 0.0f ? 1.0f / len : 0.0f;
    S.rays.u[0][r] = d0 * il; S.rays.u[1][r] = d1 * il; S.rays.u[2][r] = d2 * il;
    if (d0 == 0.0f || d1 == 0.0f || d2 == 0.0f) atomicOr(&S.flag, 1);
  }
  __syncthreads();
  const bool degenerate = S.flag != 0;
  const int nx = g.gn[0], nz = g.gn[2], ny = g.gn[1];
  bool full_scan = keep_zeros || degenerate || nx > MAX_SLABS;

  int nmarked = 0;
  if (!full_scan) {
    // per-plane windows = bounding boxes of the rays' rectangles
    for (int i = tid; i < nx; i += GEN_BLOCK) {
      S.wy0[i] = INT_MAX; S.wy1[i] = -1; S.wz0[i] = INT_MAX; S.wz1[i] = -1;
    }
    __syncthreads();
    for (int idx = tid; idx < nrays * nx; idx += GEN_BLOCK) {
      const int r = idx / nx, i = idx % nx;
      int y0, y1, z0, z1;
      if (slab_rect(g, S.rays, r, i, y0, y1, z0, z1)) {
        atomicMin(&S.wy0[i], y0); atomicMax(&S.wy1[i], y1);
        atomicMin(&S.wz0[i], z0); atomicMax(&S.wz1[i], z1);
      }
    }
    __syncthreads();
    // window areas -> bit offsets (exclusive scan over planes)
    int run = 0;
    for (int base = 0; base < nx; base += GEN_BLOCK) {
      const int i = base + tid;
      int area = 0;
      if (i < nx && S.wy1[i] >= S.wy0[i] && S.wz1[i] >= S.wz0[i])
        area = (S.wy1[i] - S.wy0[i] + 1) * (S.wz1[i] - S.wz0[i] + 1);
      int tot;
      const int ex = block_excl_sum<GEN_BLOCK>(area, S.wave, &tot);
      if (i < nx) S.off[i] = run + ex;
      run += tot;
    }
    if (tid == 0) S.off[nx] = run;
    __syncthreads();
    const int nbits = S.off[nx];
    if (nbits > BITMAP_WORDS * 32) {
      full_scan = true;
    } else {
      const int nwords = (nbits + 31) / 32;
      for (int w = tid; w < nwords; w += GEN_BLOCK) S.bits[w] = 0u;
      __syncthreads();
      for (int idx = tid; idx < nrays * nx; idx += GEN_BLOCK) {
        const int r = idx / nx, i = idx % nx;
        int y0, y1, z0, z1;
        if (!slab_rect(g, S.rays, r, i, y0, y1, z0, z1)) continue;
        const int wz = S.wz1[i] - S.wz0[i] + 1;
        for (int iy = y0; iy <= y1; iy++) {
          const int b0 = S.off[i] + (iy - S.wy0[i]) * wz - S.wz0[i];
          for (int iz = z0; iz <= z1; iz++) {
            const int b = b0 + iz;
            atomicOr(&S.bits[b >> 5], 1u << (b & 31));
          }
        }
      }
      __syncthreads();
      // enumerate marked voxels in bit order = ascending voxel id
      int pos = 0;
      for (int base = 0; base < nwords; base += GEN_BLOCK) {
        const int w = base + tid;
        uint32_t word = (w < nwords) ? S.bits[w] : 0u;
        int tot;
        const int ex = block_excl_sum<GEN_BLOCK>(__popc(word), S.wave, &tot);
        int k = pos + ex;
        while (word) {
          const int j = __ffs(word) - 1;
          word &= word - 1;
          const int b = w * 32 + j;
          // plane of bit b: last i with off[i] <= b
          int lo = 0, hi = nx - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (S.off[mid] <= b) lo = mid; else hi = mid - 1;
          }
          const int q = b - S.off[lo];
          const int wz = S.wz1[lo] - S.wz0[lo] + 1;
          const int iy = S.wy0[lo] + q / wz, iz = S.wz0[lo] + q % wz;
          if (k < rowcap) ocol[k] = (lo * ny + iy) * nz + iz;
          k++;
        }
        pos += tot;
      }
      nmarked = pos;
      if (nmarked > rowcap) {
        if (tid == 0) row_cnt[blockIdx.x] = -nmarked;
        return;   // uniform
      }
      __syncthreads();   // marked list visible to the whole workgroup
    }
  }

  float p0[3], p1[3];
  pixel_centres(g, ci, p0, p1);
  int kept = 0;
  if (!full_scan) {
    for (int base = 0; base < nmarked; base += GEN_BLOCK) {
      const int k = base + tid;
      const int v = (k < nmarked) ? ocol[k] : -1;
      __syncthreads();   // every read of this block of the list precedes the writes
      float a = 0.0f;
      bool keep = false;
      if (v >= 0 && is_candidate(g, p0, p1, v)) {
        a = row_value(g, S.rays, nrays, false, v);
        keep = a != 0.0f;
      }
      int tot;
      const int r = block_rank<GEN_BLOCK>(keep, S.wave, &tot);
      if (keep) { ocol[kept + r] = v; oval[kept + r] = a; }
      kept += tot;
    }
  } else {
    bool overflow = false;
    for (int base = 0; base < g.V; base += GEN_BLOCK) {
      const int v = base + tid;
      float a = 0.0f;
      bool keep = false;
      if (v < g.V && is_candidate(g, p0, p1, v)) {
        a = row_value(g, S.rays, nrays, degenerate, v);
        keep = keep_zeros || a != 0.0f;
      }
      int tot;
      const int r = block_rank<GEN_BLOCK>(keep, S.wave, &tot);
      if (keep && kept + r < rowcap) { ocol[kept + r] = v; oval[kept + r] = a; }
      kept += tot;
    }
    overflow = kept > rowcap;
    if (overflow) {
      if (tid == 0) row_cnt[blockIdx.x] = -kept;
      return;
    }
  }
  if (tid == 0) row_cnt[blockIdx.x] = kept;
}

// Copy each row's entries from the per-row scratch into the packed arrays.
__global__ __launch_bounds__(256) void k_rows_pack(const int* __restrict__ scol,
                                                  const float* __restrict__ sval, int rowcap,
                                                  const int* __restrict__ cnt,
                                                  const int64_t* __restrict__ off,
                                                  int64_t base, int* __restrict__ col,
                                                  float* __restrict__ val,
                                                  int64_t* __restrict__ rowptr,
                                                  int* __restrict__ row_of_entry,
                                                  int row_id) {
  const int r = blockIdx.x;
  const int n = cnt[r];
  const int64_t o = base + off[r];
  const int* sc = scol + (size_t)r * rowcap;
  const float* sv = sval + (size_t)r * rowcap;
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    col[o + k] = sc[k];
    val[o + k] = sv[k];
    if (row_of_entry) row_of_entry[o + k] = row_id;
  }
  if (rowptr && threadIdx.x == 0) rowptr[r] = o;
}

// Same, COO flavour: row id of every entry = the channel id of its row.
__global__ __launch_bounds__(256) void k_rows_pack_coo(const int* __restrict__ scol,
                                                      const float* __restrict__ sval,
                                                      int rowcap, const int* __restrict__ cnt,
                                                      const int64_t* __restrict__ off,
                                                      unsigned long long* __restrict__ dest,
                                                      const int* __restrict__ list, int row0,
                                                      int* __restrict__ col,
                                                      float* __restrict__ val,
                                                      int* __restrict__ rowid) {
  const int r = blockIdx.x;
  const int n = cnt[r];
  const long long o = (long long)(*dest) + off[r];
  const int* sc = scol + (size_t)r * rowcap;
  const float* sv = sval + (size_t)r * rowcap;
  const int cnl = list[row0 + r];
  for (int k = threadIdx.x; k < n; k += blockDim.x) {
    col[o + k] = sc[k];
    val[o + k] = sv[k];
    rowid[o + k] = cnl;
  }
}

__global__ void k_add_u64(unsigned long long* dest, const int64_t* total) {
  if (threadIdx.x == 0) *dest += (unsigned long long)(*total);
}

__global__ void k_set_i64(int64_t* p, int64_t v) {
  if (threadIdx.x == 0) *p = v;
}

// ------------------------------------------------------------------ row driver
// Generates rows [0, m) of `list` chunk by chunk into per-row scratch and hands
// each packed chunk to `sink(r0, nrows, cnt_dev, off_dev, chunk_nnz, scol,
// sval, rowcap)`.  Retries a chunk with a larger per-row capacity when a row
// overflows.
template <typename Sink>
static int generate_rows(epet_ctx* ctx, const Geo& g, const int* list, int m,
                         int keep_zeros, Sink&& sink) {
  const int V = g.V;
  long long rowcap = keep_zeros ? V : std::min<long long>(V, 4096);
  const size_t budget = (size_t)4 << 30;   // per-row scratch budget
  int* cnt = nullptr;
  int64_t* off = nullptr;
  int r0 = 0;
  int maxrows = 0;
  int rc = EPET_OK;
  while (r0 < m) {
    long long rows = (long long)(budget / ((size_t)rowcap * 8));
    rows = std::max<long long>(1, std::min<long long>(rows, 1 << 17));
    rows = std::min<long long>(rows, m - r0);
    if (rows > maxrows) {
      if (cnt) (void)hipFreeAsync(cnt, ctx->stream);
      if (off) (void)hipFreeAsync(off, ctx->stream);
      cnt = nullptr;
      off = nullptr;
      if (hipMallocAsync((void**)&cnt, sizeof(int) * rows, ctx->stream) != hipSuccess ||
          hipMallocAsync((void**)&off, sizeof(int64_t) * (rows + 1), ctx->stream) != hipSuccess) {
        (void)hipGetLastError();
        rc = fail(EPET_ERR_ALLOC, __FILE__, __LINE__, "row counts");
        break;
      }
      maxrows = (int)rows;
    }
    void* scr = nullptr;
    rc = ctx_scratch(ctx, (size_t)rows * rowcap * 8, &scr);
    if (rc != EPET_OK) break;
    int* scol = (int*)scr;
    float* sval = (float*)((int*)scr + (size_t)rows * rowcap);
    k_sm_rows<<<(unsigned)rows, GEN_BLOCK, 0, ctx->stream>>>(g, list, r0, (int)rowcap,
                                                             keep_zeros, scol, sval, cnt);
    if (hipGetLastError() != hipSuccess) { rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "k_sm_rows launch"); break; }
    // overflow check: minimum of the counts
    std::vector<int> hc(rows);
    if (hipMemcpyAsync(hc.data(), cnt, sizeof(int) * rows, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      rc = fail(EPET_ERR_HIP, __FILE__, __LINE__, "row count readback");
      break;
    }
    int need = 0;
    for (int v : hc) need = std::max(need, -v);
    if (need > 0) {
      rowcap = std::min<long long>(V, std::max<long long>(rowcap * 2, (need + 1023) / 1024 * 1024));
      continue;   // redo this chunk
    }
    long long chunk_nnz = 0;
    rc = exclusive_scan_i32_to_i64(ctx, cnt, off, rows, &chunk_nnz);
    if (rc != EPET_OK) break;
    rc = sink(r0, (int)rows, cnt, off, chunk_nnz, scol, sval, (int)rowcap);
    if (rc != EPET_OK) break;
    r0 += (int)rows;
  }
  if (cnt) (void)hipFreeAsync(cnt, ctx->stream);
  if (off) (void)hipFreeAsync(off, ctx->stream);
  return rc;
}

int matrix_alloc(epet_matrix** out, int m, int n, long long nnz, int value_type);


// ------------------------------------------------------------------ Siddon chords
// getChords (ChordsCalc_lowlevel.hpp:173-357) for a batch of rays, one thread
// per ray, in the reference's single-precision operation order (this file is
// built without contraction and with IEEE division and square root), with the
// reference's quirks: axes the ray does not cross still take part in the
// crossing comparison (:312), aDimnext starts at aDimmin + aDimup (:242), and a
// chord past the exit plane can be written (:292).  Ray r's chords go to
// chords[r nChords + k], its voxel (ix, iy, iz) to voxelIds[3 (r nChords + k)];
// counts[r] = chords written, or -1 when nChords is too small (a check the
// reference lacks).  The loop makes progress on every pass (one plane is
// crossed and alpha grows by aDimup > 0) and is capped at nChords + 1 passes.
__device__ __forceinline__ void masked_extreme3(float* out, int* good, const float a[3],
                                                const int b[3], bool want_max) {
  int g = 0;
  float v = 0.0f;
  for (int d = 0; d < 3; d++) {
    if (!b[d]) continue;
    if (!g) {
      v = a[d];
      g = 1;
    } else {
      v = want_max ? fmaxf(v, a[d]) : fminf(v, a[d]);
    }
  }
  *out = v;
  *good = g;
}

__global__ __launch_bounds__(256) void k_get_chords(const float* __restrict__ rays, int nrays,
                                                    Geo g, int nChords,
                                                    float* __restrict__ chords,
                                                    int* __restrict__ voxelIds,
                                                    int* __restrict__ counts) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrays) return;
  float ray[6];
  for (int i = 0; i < 6; i++) ray[i] = rays[r * 6 + i];
  const float* O = g.go;
  const float* D = g.gd;
  const int* N = g.gn;
  float* ch = chords + r * nChords;
  int* vid = voxelIds + r * nChords * 3;
  float admin[3], admax[3];
  int crosses[3];
  for (int d = 0; d < 3; d++) {   // getAlphaDimmin/max :484-533, alpha_from_id :421-434
    float divisor = ray[d + 3] - ray[d];
    divisor += (float)(divisor == 0.0f);
    const float a0 = (O[d] + 0.0f * D[d] - ray[d]) / divisor;
    const float a1 = (O[d] + (float)N[d] * D[d] - ray[d]) / divisor;
    admin[d] = fminf(a0, a1);
    admax[d] = fmaxf(a0, a1);
  }
  int goodPar = 1;                 // getCrossesPlanes :373-399
  for (int d = 0; d < 3; d++) {
    crosses[d] = (ray[d + 3] - ray[d] != 0.0f);
    goodPar &= (crosses[d] ||
                (!crosses[d] && ray[d] >= O[d] && ray[d] < O[d] + D[d] * (float)N[d]));
  }
  for (int d = 0; d < 3; d++) crosses[d] &= goodPar;
  float aMin, aMax;
  int aMinGood, aMaxGood;
  masked_extreme3(&aMin, &aMinGood, admin, crosses, true);
  masked_extreme3(&aMax, &aMaxGood, admax, crosses, false);
  aMinGood &= ((double)aMin >= 0. && (double)aMin <= 1.);   // :208-209
  aMaxGood &= ((double)aMax >= 0. && (double)aMax <= 1.);
  if (aMin > aMax || !aMinGood || !aMaxGood) {              // :218
    counts[r] = 0;
    return;
  }
  float aup[3], anext[3];
  int idup[3], id[3];
  for (int d = 0; d < 3; d++) {   // getAlphaDimup :612-629, getIdDimup :642-654
    const float dv = ray[d + 3] - ray[d];
    float divisor = sqrtf(dv * dv);
    divisor += (float)(int)(divisor == 0.0f);
    aup[d] = D[d] / divisor;
    idup[d] = (int)(ray[d] < ray[d + 3]) - (int)(ray[d] >= ray[d + 3]);
    anext[d] = admin[d] + aup[d];
  }
  float aNext;
  int aNextExists;
  masked_extreme3(&aNext, &aNextExists, anext, crosses, false);
  for (int d = 0; d < 3; d++)     // phi_from_alpha :453-465
    id[d] = (int)floorf((ray[d] + (0.5f * (aMin + aNext)) * (ray[d + 3] - ray[d]) - O[d]) / D[d]);
  float aCurr = aMin;
  const float dx = ray[3] - ray[0], dy = ray[4] - ray[1], dz = ray[5] - ray[2];
  const float length = sqrtf((dx * dx + dy * dy) + dz * dz);   // getLength :588-599
  int k = 0;
  for (int pass = 0; aCurr < aMax && pass <= nChords; pass++) {   // :292-356
    masked_extreme3(&aNext, &aNextExists, anext, crosses, false);
    for (int d = 0; d < 3; d++) {
      if (anext[d] == aNext) {
        if (k >= nChords) {
          counts[r] = -1;
          return;
        }
        ch[k] = 1.0f * (anext[d] - aCurr) * length;
        for (int w = 0; w < 3; w++) vid[3 * k + w] = id[w];
        k++;
        aCurr = anext[d];
        anext[d] += aup[d];
        id[d] += idup[d];
      }
    }
  }
  counts[r] = aCurr < aMax ? -1 : k;
}

}  // namespace epet

using namespace epet;

// ====================================================================== C-ABI
namespace epet {
// channel id of row i: list[i], or i without a list
__global__ __launch_bounds__(256) void k_row_ids(const int* __restrict__ list, int m,
                                                 int* __restrict__ ids) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) ids[i] = list ? list[i] : (int)i;
}
}  // namespace epet

extern "C" {

int epet_rays(epet_ctx* ctx, const int* cnl_dev, int n, float* rays_dev) {
  EPET_CHECK(ctx && (n == 0 || (cnl_dev && rays_dev)), EPET_ERR_ARG, "null argument");
  EPET_CHECK(n >= 0, EPET_ERR_ARG, "negative count");
  Geo g;
  EPET_TRY(ctx_geo(ctx, &g));
  if (n == 0) return EPET_OK;
  const long long tot = (long long)n * g.nrays;
  k_rays<<<grid_for(tot, 256, 1ll << 31), 256, 0, ctx->stream>>>(g, cnl_dev, n, rays_dev);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

int epet_getWorkqueue(epet_ctx* ctx, const int* list_dev, int list_size, int* listId,
                      int* vxlId, long long cap, int* out_cnl_dev, int* out_vxl_dev,
                      long long out_capacity, long long* n_found) {
  EPET_CHECK(ctx && listId && vxlId && n_found, EPET_ERR_ARG, "null argument");
  EPET_CHECK(list_size >= 0 && (list_size == 0 || list_dev), EPET_ERR_ARG, "bad list");
  Geo g;
  EPET_TRY(ctx_geo(ctx, &g));
  *n_found = 0;
  const int li = *listId, vi = *vxlId;
  if (li >= list_size || cap == 0) return EPET_OK;
  EPET_CHECK(li >= 0 && vi >= 0 && vi < g.V, EPET_ERR_ARG,
             "Invalid voxel index encountered during function 'getWorkqueueEntry(...)': " +
                 std::to_string(vi));
  const int rows = list_size - li;
  StreamTemps tmp(ctx->stream);
  int* counts = nullptr;
  int64_t* offs = nullptr;
  EPET_HIP(tmp.alloc(&counts, sizeof(int) * rows));
  EPET_HIP(tmp.alloc(&offs, sizeof(int64_t) * (rows + 1)));
  k_wq_count<<<rows, GEN_BLOCK, 0, ctx->stream>>>(g, list_dev, li, vi, counts);
  EPET_HIP(hipGetLastError());
  long long total = 0;
  int rc = exclusive_scan_i32_to_i64(ctx, counts, offs, rows, &total);
  if (rc != EPET_OK) return rc;
  const long long want = cap < 0 ? total : std::min(cap, total);
  if (!out_cnl_dev && !out_vxl_dev && out_capacity == 0) {
    // count-only query: the number of pairs a call with this cap would emit;
    // the cursor is left unchanged
    EPET_HIP(hipStreamSynchronize(ctx->stream));
    *n_found = want;
    return EPET_OK;
  }
  if (want > out_capacity || (want > 0 && (!out_cnl_dev || !out_vxl_dev))) {
    EPET_FAIL(EPET_ERR_CAPACITY, "workqueue needs " + std::to_string(want) +
                                     " entries, output holds " + std::to_string(out_capacity));
  }
  if (want > 0) {
    k_wq_emit<<<rows, GEN_BLOCK, 0, ctx->stream>>>(g, list_dev, li, vi, offs, want, out_cnl_dev,
                                                   out_vxl_dev);
    EPET_HIP(hipGetLastError());
  }
  // cursor (getWorkqueue.hpp:83-108, :143-169): just past the last emitted
  // pair when the cap was reached, else the end of the list.
  if (cap >= 0 && want == cap) {
    std::vector<int64_t> ho(rows);
    EPET_HIP(hipMemcpyAsync(ho.data(), offs, sizeof(int64_t) * rows, hipMemcpyDeviceToHost,
                            ctx->stream));
    int lastv = 0;
    EPET_HIP(hipMemcpyAsync(&lastv, out_vxl_dev + (want - 1), sizeof(int),
                            hipMemcpyDeviceToHost, ctx->stream));
    EPET_HIP(hipStreamSynchronize(ctx->stream));
    // row holding entry want-1: last r with ho[r] <= want-1
    const int r = (int)(std::upper_bound(ho.begin(), ho.end(), (int64_t)(want - 1)) - ho.begin()) - 1;
    if (lastv < g.V - 1) { *listId = li + r; *vxlId = lastv + 1; }
    else { *listId = li + r + 1; *vxlId = 0; }
  } else {
    *listId = list_size;
    *vxlId = 0;
    EPET_HIP(hipStreamSynchronize(ctx->stream));
  }
  *n_found = want;
  return EPET_OK;
}

int epet_getSystemMatrixFromWorkqueue(epet_ctx* ctx, const int* cnl_dev, const int* vxl_dev,
                                      float* val_dev, long long wqLength, int nrays) {
  EPET_CHECK(ctx && (wqLength == 0 || (cnl_dev && vxl_dev && val_dev)), EPET_ERR_ARG,
             "null argument");
  EPET_CHECK(wqLength >= 0, EPET_ERR_ARG, "negative workqueue length");
  EPET_CHECK(nrays >= 1 && nrays <= EPET_MAX_RAYS, EPET_ERR_UNSUPPORTED, "bad ray count");
  Geo g;
  EPET_TRY(ctx_geo(ctx, &g));
  g.nrays = nrays;
  if (wqLength == 0) return EPET_OK;
  k_sm_from_wq<<<grid_for(wqLength, 256, 1ll << 31), 256, 0, ctx->stream>>>(g, cnl_dev, vxl_dev,
                                                                           val_dev, wqLength);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

int epet_getSystemMatrix(epet_ctx* ctx, float* sme_dev, int* vxlId_dev, int* cnlId_dev,
                         const int* ml_dev, uint32_t mlSize, unsigned long long* truckDest_dev,
                         long long capacity, int keep_zeros) {
  EPET_CHECK(ctx && truckDest_dev, EPET_ERR_ARG, "null argument");
  EPET_CHECK(mlSize == 0 || (ml_dev && sme_dev && vxlId_dev && cnlId_dev), EPET_ERR_ARG,
             "null argument");
  Geo g;
  EPET_TRY(ctx_geo(ctx, &g));
  if (mlSize == 0) return EPET_OK;
  unsigned long long start = 0;
  EPET_TRY(read_pinned(ctx, &start, truckDest_dev, sizeof(start)));
  long long written = (long long)start;
  int64_t* tot_dev = nullptr;
  EPET_HIP(hipMallocAsync((void**)&tot_dev, sizeof(int64_t), ctx->stream));
  auto sink = [&](int r0, int rows, int* cnt, int64_t* off, long long chunk_nnz, int* scol,
                  float* sval, int rowcap) -> int {
    if (written + chunk_nnz > capacity)
      EPET_FAIL(EPET_ERR_CAPACITY, "system matrix needs more than " + std::to_string(capacity) +
                                       " entries");
    k_rows_pack_coo<<<rows, 256, 0, ctx->stream>>>(scol, sval, rowcap, cnt, off, truckDest_dev,
                                                   ml_dev, r0, vxlId_dev, sme_dev, cnlId_dev);
    EPET_HIP(hipGetLastError());
    k_set_i64<<<1, 64, 0, ctx->stream>>>(tot_dev, (int64_t)chunk_nnz);
    k_add_u64<<<1, 64, 0, ctx->stream>>>(truckDest_dev, tot_dev);
    EPET_HIP(hipGetLastError());
    written += chunk_nnz;
    return EPET_OK;
  };
  int rc = generate_rows(ctx, g, ml_dev, (int)mlSize, keep_zeros ? 1 : 0, sink);
  (void)hipFreeAsync(tot_dev, ctx->stream);
  if (rc == EPET_OK) EPET_HIP(hipStreamSynchronize(ctx->stream));
  return rc;
}

int epet_systemMatrixCalculation(epet_ctx* ctx, int* ecsrPtr_dev, int* vxlId_dev,
                                 float* val_dev, unsigned long long* nnz_dev, int* cnlId_dev,
                                 int* csrPtr_dev, const int* yRowId_dev, uint32_t nY,
                                 int nchannels, long long capacity, int keep_zeros) {
  // wrappers.hpp:172-247: generation, sort, COO->CSR over all channels,
  // CSR->ECSR for the nY listed rows.  nnz_dev accumulates (initial value +
  // entries), as nnz_devi does.
  EPET_CHECK(ctx && ecsrPtr_dev && nnz_dev && csrPtr_dev, EPET_ERR_ARG, "null argument");
  EPET_CHECK(nchannels > 0, EPET_ERR_ARG, "channel count must be positive");
  EPET_TRY(epet_getSystemMatrix(ctx, val_dev, vxlId_dev, cnlId_dev, yRowId_dev, nY, nnz_dev,
                                capacity, keep_zeros));
  unsigned long long nnz = 0;
  EPET_TRY(read_pinned(ctx, &nnz, nnz_dev, sizeof(nnz)));
  EPET_TRY(epet_cooSort(ctx, val_dev, cnlId_dev, vxlId_dev, (long long)nnz));
  EPET_TRY(epet_convertCoo2Csr(ctx, csrPtr_dev, cnlId_dev, (long long)nnz, nchannels));
  EPET_TRY(epet_convertCsr2Ecsr(ctx, ecsrPtr_dev, yRowId_dev, (int)nY, csrPtr_dev, nchannels));
  EPET_HIP(hipStreamSynchronize(ctx->stream));
  return EPET_OK;
}

int epet_sm_build(epet_ctx* ctx, const int* list_dev, int m, epet_matrix** out) {
  EPET_CHECK(ctx && out, EPET_ERR_ARG, "null argument");
  EPET_CHECK(m >= 0, EPET_ERR_ARG, "negative row count");
  Geo g;
  EPET_TRY(ctx_geo(ctx, &g));
  EPET_CHECK(list_dev != nullptr || m <= g.nch, EPET_ERR_ARG, "more rows than channels");
  // packed output grows geometrically; rows are appended in order
  long long capacity = 0, used = 0;
  int* col = nullptr;
  float* val = nullptr;
  int64_t* rowptr = nullptr;
  EPET_HIP(hipMalloc((void**)&rowptr, sizeof(int64_t) * ((size_t)m + 1)));
  auto grow = [&](long long need) -> int {
    if (need <= capacity) return EPET_OK;
    long long cap2 = std::max<long long>(need, capacity + capacity / 4);
    int* c2 = nullptr;
    float* v2 = nullptr;
    if (hipMalloc((void**)&c2, sizeof(int) * (size_t)std::max(1ll, cap2)) != hipSuccess ||
        hipMalloc((void**)&v2, sizeof(float) * (size_t)std::max(1ll, cap2)) != hipSuccess) {
      (void)hipGetLastError();
      if (c2) (void)hipFree(c2);
      EPET_FAIL(EPET_ERR_ALLOC, "system matrix allocation of " + std::to_string(cap2) +
                                    " entries failed");
    }
    if ((used > 0 &&
         (hipMemcpyAsync(c2, col, sizeof(int) * used, hipMemcpyDeviceToDevice, ctx->stream) !=
              hipSuccess ||
          hipMemcpyAsync(v2, val, sizeof(float) * used, hipMemcpyDeviceToDevice, ctx->stream) !=
              hipSuccess)) ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      (void)hipFree(c2);
      (void)hipFree(v2);
      EPET_FAIL(EPET_ERR_HIP, "system matrix: growing the arrays");
    }
    if (col) (void)hipFree(col);
    if (val) (void)hipFree(val);
    col = c2; val = v2; capacity = cap2;
    return EPET_OK;
  };
  auto sink = [&](int r0, int rows, int* cnt, int64_t* off, long long chunk_nnz, int* scol,
                  float* sval, int rowcap) -> int {
    if (capacity == 0) {
      // first chunk sets the estimate for the whole matrix (+12%)
      long long est = (long long)((double)chunk_nnz / rows * m * 1.12) + 1024;
      EPET_TRY(grow(std::max(est, chunk_nnz)));
    } else if (used + chunk_nnz > capacity) {
      EPET_TRY(grow(used + chunk_nnz + (long long)((double)chunk_nnz / rows * (m - r0 - rows) * 1.12)));
    }
    k_rows_pack<<<rows, 256, 0, ctx->stream>>>(scol, sval, rowcap, cnt, off, used, col, val,
                                               rowptr + r0, nullptr, 0);
    EPET_HIP(hipGetLastError());
    used += chunk_nnz;
    return EPET_OK;
  };
  int rc = generate_rows(ctx, g, list_dev, m, 0, sink);
  if (rc == EPET_OK) {
    k_set_i64<<<1, 64, 0, ctx->stream>>>(rowptr + m, (int64_t)used);
    rc = (hipGetLastError() == hipSuccess && hipStreamSynchronize(ctx->stream) == hipSuccess)
             ? EPET_OK
             : fail(EPET_ERR_HIP, __FILE__, __LINE__, "finalising system matrix");
  }
  if (rc != EPET_OK) {
    (void)hipStreamSynchronize(ctx->stream);
    if (col) (void)hipFree(col);
    if (val) (void)hipFree(val);
    (void)hipFree(rowptr);
    return rc;
  }
  // channel ids of the rows (the list, or 0..m-1) and the geometry they refer to
  int* ids = nullptr;
  if (hipMalloc((void**)&ids, sizeof(int) * (size_t)std::max(1, m)) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(col);
    (void)hipFree(val);
    (void)hipFree(rowptr);
    EPET_FAIL(EPET_ERR_ALLOC, "row id allocation failed");
  }
  epet_matrix* A = new epet_matrix();
  A->m = m;
  A->n = g.V;
  A->nnz = used;
  A->value_type = EPET_VAL_F32;
  A->rowptr = rowptr;
  A->col = col;
  A->val = val;
  A->row_ids = ids;
  A->mg.na = g.na;
  A->mg.n0z = g.n0z;
  A->mg.n0y = g.n0y;
  A->mg.n1z = g.n1z;
  A->mg.n1y = g.n1y;
  A->mg.nx = g.gn[0];
  A->mg.ny = g.gn[1];
  A->mg.nz = g.gn[2];
  A->device = ctx->device;
  MatrixGuard keep(A);
  if (m > 0) {
    k_row_ids<<<grid_for(m, 256, 1ll << 31), 256, 0, ctx->stream>>>(list_dev, m, ids);
    EPET_HIP(hipGetLastError());
    EPET_HIP(hipStreamSynchronize(ctx->stream));
  }
  *out = keep.release();
  return EPET_OK;
}


int epet_getChords(epet_ctx* ctx, const float* rays_dev, int nrays, int nChords,
                   float* chords_dev, int* voxelIds_dev, int* counts_dev) {
  EPET_CHECK(ctx && (nrays == 0 || (rays_dev && chords_dev && voxelIds_dev && counts_dev)),
             EPET_ERR_ARG, "null argument");
  EPET_CHECK(nrays >= 0 && nChords >= 0, EPET_ERR_ARG, "negative size");
  Geo g;
  EPET_TRY(ctx_geo(ctx, &g));
  if (nrays > 0)
    k_get_chords<<<grid_for(nrays, 256, 1ll << 31), 256, 0, ctx->stream>>>(
        rays_dev, nrays, g, nChords, chords_dev, voxelIds_dev, counts_dev);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

}  // extern "C"
