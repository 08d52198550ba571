// epet_internal.hpp -- shared host/device definitions of libepet.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/epet.h"

namespace epet {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const char* file, int line, const std::string& msg);

#define EPET_FAIL(code, msg) return ::epet::fail((code), __FILE__, __LINE__, (msg))
#define EPET_HIP(call)                                                        \
  do {                                                                        \
    hipError_t _e = (call);                                                   \
    if (_e != hipSuccess)                                                     \
      return ::epet::fail(EPET_ERR_HIP, __FILE__, __LINE__,                   \
                          std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)
#define EPET_CHECK(cond, code, msg) \
  do { if (!(cond)) EPET_FAIL((code), (msg)); } while (0)
#define EPET_TRY(call) \
  do { int _r = (call); if (_r != EPET_OK) return _r; } while (0)

// ---------------------------------------------------------------- temporaries
// Stream-ordered temporary device buffers, freed on the stream when the guard
// leaves scope, so that every return path (error paths included) releases
// them after the work enqueued before it.
class StreamTemps {
 public:
  explicit StreamTemps(hipStream_t st) : st_(st) {}
  ~StreamTemps() {
    for (int i = 0; i < n_; i++) (void)hipFreeAsync(p_[i], st_);
  }
  StreamTemps(const StreamTemps&) = delete;
  StreamTemps& operator=(const StreamTemps&) = delete;
  template <typename T>
  hipError_t alloc(T** out, size_t bytes) {
    *out = nullptr;
    if (n_ == kMax) return hipErrorOutOfMemory;
    const hipError_t e = hipMallocAsync(reinterpret_cast<void**>(out), bytes, st_);
    if (e == hipSuccess) p_[n_++] = *out;
    return e;
  }

 private:
  static constexpr int kMax = 8;
  hipStream_t st_;
  void* p_[kMax] = {};
  int n_ = 0;
};

// A matrix under construction, destroyed unless released to the caller.
class MatrixGuard {
 public:
  explicit MatrixGuard(epet_matrix* A) : A_(A) {}
  ~MatrixGuard() {
    if (A_) (void)epet_matrix_destroy(A_);
  }
  MatrixGuard(const MatrixGuard&) = delete;
  MatrixGuard& operator=(const MatrixGuard&) = delete;
  epet_matrix* release() {
    epet_matrix* A = A_;
    A_ = nullptr;
    return A;
  }

 private:
  epet_matrix* A_;
};

// ---------------------------------------------------------------- geometry
// Everything a generation kernel needs, passed by value in the kernel-argument
// segment (read-only, scalar-cache resident: the AMD home of per-launch
// uniform data that CUDA keeps in __constant__ banks).
struct Geo {
  // measurement setup (MeasurementSetup.hpp:129-139 fields)
  float pos0x, pos1x;
  int na;
  float da;
  int n0z, n0y, n1z, n1y;
  float segx, segy, segz;
  // voxel grid (VoxelGrid.hpp:88-91)
  float go[3], gd[3];
  int gn[3];
  // derived
  int V;                 // voxels
  int nch;               // channels
  int P1, P2, P3, P4;    // channel-id radix products n1y, n1y*n1z, .., all four
  float sumRadii;        // getWorkqueue.hpp:54-56
  float halfDiag2;       // (voxel half diagonal)^2 for the conservative ray test
  int nrays;
  uint32_t key0, key1;   // Philox key = seed
  float sin_[EPET_MAX_ANGLES];
  float cos_[EPET_MAX_ANGLES];
};

}  // namespace epet

// ---------------------------------------------------------------- context
// Path choices of a context, read once from the environment when it is created
// (for comparisons and measurements; the defaults are the fast paths):
//   EPET_LDS=0         single-frame products without LDS-tiled copies
//   EPET_LDS_MAP=1     LDS-tiled copies of grid matrices in y-major voxel order
//   EPET_BP=atomic|tiled|rows   single-frame back-projection variant
//   EPET_BATCHED=gather         multi-frame MLEM with the gather kernels
enum { BP_DEFAULT = 0, BP_ATOMIC = 1, BP_TILED = 2, BP_ROWS = 3 };
struct EpetOpts {
  bool lds = true;
  bool lds_map = false;
  int bp = BP_DEFAULT;
  bool batched_gather = false;
};

struct epet_ctx {
  EpetOpts opt;
  int device = 0;
  hipStream_t stream = nullptr;
  bool have_geo = false;
  epet::Geo geo{};
  int nrays = 8;
  uint64_t seed = 0;
  // small pinned host buffer for synchronous scalar reads
  void* pinned = nullptr;
  // reusable device scratch
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  int num_cus = 256;
};

// Column order of a tile-major copy.  nx = 0: the matrix's own column order.
// Otherwise the columns are voxels (ix ny + iy) nz + iz of an nx x ny x nz
// grid, renumbered y-major, (iy nx + ix) nz + iz, so that tiles are slabs of
// y-planes: a channel's line keeps its y within a few planes, so its row
// meets few tiles.
struct ColMap {
  int nx = 0, ny = 0, nz = 0;
};
__host__ __device__ inline long long colmap_fwd(long long c, const ColMap& mp) {
  if (mp.nx == 0) return c;
  const long long iz = c % mp.nz, q = c / mp.nz;
  const long long iy = q % mp.ny, ix = q / mp.ny;
  return (iy * mp.nx + ix) * mp.nz + iz;
}
__host__ __device__ inline long long colmap_inv(long long pc, const ColMap& mp) {
  if (mp.nx == 0) return pc;
  const long long iz = pc % mp.nz, q = pc / mp.nz;
  const long long ix = q % mp.nx, iy = q / mp.nx;
  return (ix * mp.ny + iy) * mp.nz + iz;
}

// Geometry a generated matrix was built on: the channel-id radices of its
// rows (MeasurementSetupLinIndex.hpp:40-44) and the voxel grid of its columns
// (z fastest).  All zero when unknown (caller-built matrices).
struct MatGeo {
  int na = 0, n0z = 0, n0y = 0, n1z = 0, n1y = 0;
  int nx = 0, ny = 0, nz = 0;
};

namespace epet { struct BlkCopy; }

struct epet_matrix {
  int m = 0, n = 0;
  long long nnz = 0;
  int value_type = EPET_VAL_F32;
  int64_t* rowptr = nullptr;   // m+1
  int* col = nullptr;          // nnz
  void* val = nullptr;         // nnz x (4 | 2) bytes
  float* rowscale = nullptr;   // m (reduced-precision storage only)
  int* row_ids = nullptr;      // m: channel id of every row, when known
  MatGeo mg;
  int device = 0;
  bool owns = true;
  // tuning chosen for this matrix (lanes per row of the row-parallel kernels)
  int lanes_fwd = 0;
  // back-projection plan (built on first use): voxel tiles of bp_tile ids,
  // bp_T tiles, per-row entry offsets of every tile boundary (m x (bp_T+1)),
  // partial images of bp_nblk row blocks
  int bp_T = 0, bp_tile = 0, bp_nblk = 0;
  uint32_t* bp_toff = nullptr;
  float* bp_partial = nullptr;
  bool bp_failed = false;
  bool cols_strict = true;   // columns strictly ascending within every row
  // transposed copy (CSR of A^T: rows = voxels, entries (row index, value)),
  // used by the back projections; per-row scales of A are folded into the
  // back-projected vector instead
  epet_matrix* tr = nullptr;
  // on a transposed copy: entry offsets (relative to the row start) of the
  // XB_BLOCKS contiguous blocks of its columns, (m x (XB_BLOCKS+1)), and the
  // per-block partial products (XB_BLOCKS x m)
  uint32_t* xb_toff = nullptr;
  float* xb_partial = nullptr;
  // LDS-tiled products: the gathered vector is cut into lds_T tiles of
  // LDS_TILE entries, and the matrix is kept a second time in tile-major
  // order: the entries of (tile t, row r) are lds_off[t m + r] .. lds_off[t m +
  // r + 1] of lds_val (stored value type) and lds_col16 (tile-local column,
  // col mod LDS_TILE); lds_partial holds the per-tile partial products (lds_T x m)
  int lds_T = 0;
  int64_t* lds_off = nullptr;
  void* lds_val = nullptr;
  uint16_t* lds_col16 = nullptr;
  float* lds_partial = nullptr;
  ColMap lds_map;
  bool lds_failed = false;
  // blocked multi-frame copy (register-accumulated products, one per padded
  // frame count in use) and its autotuned launch shape
  epet::BlkCopy* blk = nullptr;
  // multi-frame kernel configuration picked by the autotuner for mf_F frames
  // (rows per wavefront R, frames per lane tile P) and the times it measured
  int mf_F = 0, mf_fwd_R = 1, mf_fwd_P = 32, mf_bwd_R = 1, mf_bwd_P = 32;
  float mf_fwd_ms = 0.0f, mf_bwd_ms = 0.0f;
  ~epet_matrix();
};

// column blocks of the blocked back projection: one per XCD of the MI355X
constexpr int XB_BLOCKS = 8;
// LDS tile of the LDS-tiled row products: 32768 floats = 128 KB of the 160 KB
// of LDS per workgroup on gfx950
constexpr int LDS_TILE = 32768;
constexpr int LDS_CHUNK = 8;        // tile-major segments are whole chunks of 8 entries



namespace epet {

void blk_release(epet_matrix* A);   // frees the blocked copies of A

// device scratch owned by the context, grown on demand (stream-ordered use)
int ctx_scratch(epet_ctx* ctx, size_t bytes, void** out);
int ctx_geo(epet_ctx* ctx, Geo* g);   // copy of the geometry with nrays/seed
int read_pinned(epet_ctx* ctx, void* dst, const void* src_dev, size_t bytes);

// ---------------------------------------------------------------- scans
// Exclusive prefix sums over n entries (device pointers), stream-ordered.
// out may alias in.  Returns the total on the host when total != nullptr
// (synchronises).
int exclusive_scan_i32_to_i64(epet_ctx* ctx, const int* in, int64_t* out, long long n,
                              long long* total);
int exclusive_scan_i64(epet_ctx* ctx, const int64_t* in, int64_t* out, long long n,
                       long long* total);
int exclusive_scan_i32(epet_ctx* ctx, const int* in, int* out, long long n,
                       long long* total);

// ---------------------------------------------------------------- multi-frame
// Buffers of one multi-frame MLEM iteration (frame-minor, F frames).
struct MfArgs {
  int F;
  const float *x, *y;
  float *e, *c;
  const int* itp;
  double* loglik;
};

// ---------------------------------------------------------------- launch helpers
inline unsigned grid_for(long long n, int block, long long max_blocks = 1 << 20) {
  long long b = (n + block - 1) / block;
  if (b < 1) b = 1;
  if (b > max_blocks) b = max_blocks;
  return (unsigned)b;
}

}  // namespace epet
