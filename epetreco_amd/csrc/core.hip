// core.hip -- errors, contexts, geometry upload, device memory, prefix sums,
// chunk helpers of libepet.so.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "epet_block.hpp"

#include "epet_internal.hpp"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

namespace epet {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const char* file, int line, const std::string& msg) {
  // file(line) prefix, as HANDLE_ERROR prints (CUDA_HandleError.hpp:17-31)
  const char* base = std::strrchr(file, '/');
  g_last_error = std::string(base ? base + 1 : file) + "(" + std::to_string(line) + "): " + msg;
  return code;
}

int ctx_scratch(epet_ctx* ctx, size_t bytes, void** out) {
  if (bytes > ctx->scratch_bytes) {
    if (ctx->scratch) {
      EPET_HIP(hipStreamSynchronize(ctx->stream));
      EPET_HIP(hipFree(ctx->scratch));
      ctx->scratch = nullptr;
      ctx->scratch_bytes = 0;
    }
    size_t want = bytes < (1u << 20) ? (1u << 20) : bytes;
    if (hipMalloc(&ctx->scratch, want) != hipSuccess) {
      (void)hipGetLastError();
      EPET_FAIL(EPET_ERR_ALLOC, "scratch allocation of " + std::to_string(want) + " bytes failed");
    }
    ctx->scratch_bytes = want;
  }
  *out = ctx->scratch;
  return EPET_OK;
}

int ctx_geo(epet_ctx* ctx, Geo* g) {
  EPET_CHECK(ctx->have_geo, EPET_ERR_STATE, "geometry not set (epet_set_geometry)");
  *g = ctx->geo;
  g->nrays = ctx->nrays;
  g->key0 = (uint32_t)ctx->seed;
  g->key1 = (uint32_t)(ctx->seed >> 32);
  return EPET_OK;
}

int read_pinned(epet_ctx* ctx, void* dst, const void* src_dev, size_t bytes) {
  EPET_CHECK(bytes <= 4096, EPET_ERR_ARG, "pinned read too large");
  EPET_HIP(hipMemcpyAsync(ctx->pinned, src_dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
  EPET_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(dst, ctx->pinned, bytes);
  return EPET_OK;
}

// ------------------------------------------------------------------ scans
// Three-phase exclusive scan: per-tile sums, scan of tile sums (recursive),
// per-tile scan plus tile offset.  TILE = 256 threads x 8 items.
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr long long SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

template <typename Tin, typename Tout>
__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_sums(const Tin* __restrict__ in, long long n,
                                                         Tout* __restrict__ sums) {
  __shared__ Tout s[SCAN_BLOCK / WAVE];
  const long long base = (long long)blockIdx.x * SCAN_TILE;
  Tout acc = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; j++) {
    const long long i = base + (long long)j * SCAN_BLOCK + threadIdx.x;
    if (i < n) acc += (Tout)in[i];
  }
  acc = wave_sum(acc);
  if (lane_id() == 0) s[wave_id()] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    Tout t = 0;
    for (int w = 0; w < SCAN_BLOCK / WAVE; w++) t += s[w];
    sums[blockIdx.x] = t;
  }
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_scan(const Tin* __restrict__ in, long long n,
                                                         const Tout* __restrict__ offs,
                                                         Tout* __restrict__ out) {
  __shared__ Tout s_wave[SCAN_BLOCK / WAVE];
  const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
  Tout v[SCAN_ITEMS];
  Tout local = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; j++) {
    const long long i = base + j;
    v[j] = (i < n) ? (Tout)in[i] : (Tout)0;
    local += v[j];
  }
  // exclusive scan of `local` over the block
  Tout x = local;
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const Tout y = __shfl_up(x, off, WAVE);
    if (lane >= off) x += y;
  }
  if (lane == WAVE - 1) s_wave[wave_id()] = x;
  __syncthreads();
  Tout wbase = 0;
  for (int w = 0; w < wave_id(); w++) wbase += s_wave[w];
  Tout run = offs[blockIdx.x] + wbase + x - local;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; j++) {
    const long long i = base + j;
    if (i < n) out[i] = run;
    run += v[j];
  }
}

template <typename T>
__global__ void k_set_scalar(T* p, T v) { if (threadIdx.x == 0) *p = v; }

template <typename T>
__global__ void k_first_tile_fix(T* sums) {
  // single tile: total = sums[0], exclusive offset of the tile = 0
  if (threadIdx.x == 0) { sums[1] = sums[0]; sums[0] = 0; }
}

template <typename Tin, typename Tout>
static int scan_impl(epet_ctx* ctx, const Tin* in, Tout* out, long long n, Tout* total_dev) {
  // out may alias in: every tile reads its inputs into registers before it
  // writes, and tiles are disjoint.
  if (n <= 0) {
    if (total_dev) {
      k_set_scalar<Tout><<<1, 64, 0, ctx->stream>>>(total_dev, (Tout)0);
      EPET_HIP(hipGetLastError());
    }
    return EPET_OK;
  }
  const long long ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  EPET_CHECK(ntiles < (1ll << 31), EPET_ERR_UNSUPPORTED, "scan too long");
  StreamTemps tmp(ctx->stream);
  Tout* sums = nullptr;
  EPET_HIP(tmp.alloc(&sums, sizeof(Tout) * (ntiles + 1)));
  k_tile_sums<Tin, Tout><<<(unsigned)ntiles, SCAN_BLOCK, 0, ctx->stream>>>(in, n, sums);
  EPET_HIP(hipGetLastError());
  // scan of the tile sums; the total lands in sums[ntiles]
  if (ntiles == 1) {
    k_first_tile_fix<Tout><<<1, 64, 0, ctx->stream>>>(sums);
    EPET_HIP(hipGetLastError());
  } else {
    Tout* total2 = sums + ntiles;
    EPET_TRY((scan_impl<Tout, Tout>(ctx, sums, sums, ntiles, total2)));
  }
  k_tile_scan<Tin, Tout><<<(unsigned)ntiles, SCAN_BLOCK, 0, ctx->stream>>>(in, n, sums, out);
  EPET_HIP(hipGetLastError());
  if (total_dev)
    EPET_HIP(hipMemcpyAsync(total_dev, sums + ntiles, sizeof(Tout), hipMemcpyDeviceToDevice,
                            ctx->stream));
  return EPET_OK;
}

template <typename Tin, typename Tout>
static int scan_total(epet_ctx* ctx, const Tin* in, Tout* out, long long n, long long* total) {
  if (!total) return scan_impl<Tin, Tout>(ctx, in, out, n, nullptr);
  Tout* tdev = nullptr;
  EPET_HIP(hipMallocAsync((void**)&tdev, sizeof(Tout), ctx->stream));
  int r = scan_impl<Tin, Tout>(ctx, in, out, n, tdev);
  Tout th = 0;
  if (r == EPET_OK) r = read_pinned(ctx, &th, tdev, sizeof(Tout));
  (void)hipFreeAsync(tdev, ctx->stream);
  *total = (long long)th;
  return r;
}

int exclusive_scan_i32_to_i64(epet_ctx* ctx, const int* in, int64_t* out, long long n,
                              long long* total) {
  return scan_total<int, int64_t>(ctx, in, out, n, total);
}
int exclusive_scan_i64(epet_ctx* ctx, const int64_t* in, int64_t* out, long long n,
                       long long* total) {
  return scan_total<int64_t, int64_t>(ctx, in, out, n, total);
}
int exclusive_scan_i32(epet_ctx* ctx, const int* in, int* out, long long n, long long* total) {
  return scan_total<int, int>(ctx, in, out, n, total);
}

}  // namespace epet

using namespace epet;

// ====================================================================== C-ABI
extern "C" {

const char* epet_last_error(void) { return g_last_error.c_str(); }

int epet_version(void) { return 1; }

int epet_ctx_create(epet_ctx** out, int device, void* stream) {
  EPET_CHECK(out != nullptr, EPET_ERR_ARG, "null output pointer");
  int ndev = 0;
  EPET_HIP(hipGetDeviceCount(&ndev));
  EPET_CHECK(device >= 0 && device < ndev, EPET_ERR_ARG,
             "device " + std::to_string(device) + " not present (" + std::to_string(ndev) +
                 " devices)");
  EPET_HIP(hipSetDevice(device));
  epet_ctx* c = new epet_ctx();
  c->device = device;
  {
    const char* e = getenv("EPET_LDS");
    c->opt.lds = !(e && e[0] == '0');
    e = getenv("EPET_LDS_MAP");
    c->opt.lds_map = e && e[0] == '1';
    e = getenv("EPET_BP");
    const std::string bp = e ? e : "";
    c->opt.bp = bp == "atomic" ? BP_ATOMIC : bp == "tiled" ? BP_TILED : bp == "rows" ? BP_ROWS
                                                                                : BP_DEFAULT;
    e = getenv("EPET_BATCHED");
    c->opt.batched_gather = e && std::string(e) == "gather";
  }
  c->stream = (hipStream_t)stream;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  if (hipHostMalloc(&c->pinned, 4096, hipHostMallocDefault) != hipSuccess) {
    delete c;
    EPET_FAIL(EPET_ERR_ALLOC, "pinned host allocation failed");
  }
  *out = c;
  return EPET_OK;
}

int epet_ctx_destroy(epet_ctx* ctx) {
  if (!ctx) return EPET_OK;
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  delete ctx;
  return EPET_OK;
}

int epet_ctx_set_stream(epet_ctx* ctx, void* stream) {
  EPET_CHECK(ctx, EPET_ERR_ARG, "null context");
  ctx->stream = (hipStream_t)stream;
  return EPET_OK;
}

int epet_ctx_synchronize(epet_ctx* ctx) {
  EPET_CHECK(ctx, EPET_ERR_ARG, "null context");
  EPET_HIP(hipStreamSynchronize(ctx->stream));
  return EPET_OK;
}

int epet_set_geometry(epet_ctx* ctx, const epet_setup* s, const epet_grid* g) {
  EPET_CHECK(ctx && s && g, EPET_ERR_ARG, "null argument");
  EPET_CHECK(s->na > 0 && s->na <= EPET_MAX_ANGLES, EPET_ERR_UNSUPPORTED,
             "number of angles must be in [1, " + std::to_string(EPET_MAX_ANGLES) + "]");
  EPET_CHECK(s->n0z > 0 && s->n0y > 0 && s->n1z > 0 && s->n1y > 0, EPET_ERR_ARG,
             "segment counts must be positive");
  EPET_CHECK(g->n[0] > 0 && g->n[1] > 0 && g->n[2] > 0, EPET_ERR_ARG,
             "grid dimensions must be positive");
  const long long V = (long long)g->n[0] * g->n[1] * g->n[2];
  const long long nch = (long long)s->na * s->n0z * s->n0y * s->n1z * s->n1y;
  EPET_CHECK(V < (1ll << 31) && nch < (1ll << 31), EPET_ERR_UNSUPPORTED,
             "grid or channel count exceeds int32 ids");
  Geo& G = ctx->geo;
  std::memset(&G, 0, sizeof(Geo));
  G.pos0x = s->pos0x; G.pos1x = s->pos1x; G.na = s->na; G.da = s->da;
  G.n0z = s->n0z; G.n0y = s->n0y; G.n1z = s->n1z; G.n1y = s->n1y;
  G.segx = s->segx; G.segy = s->segy; G.segz = s->segz;
  for (int d = 0; d < 3; d++) { G.go[d] = g->o[d]; G.gd[d] = g->d[d]; G.gn[d] = g->n[d]; }
  G.V = (int)V;
  G.nch = (int)nch;
  G.P1 = s->n1y;
  G.P2 = s->n1y * s->n1z;
  G.P3 = s->n1y * s->n1z * s->n0y;
  G.P4 = s->n1y * s->n1z * s->n0y * s->n0z;
  {
    // getWorkqueue.hpp:54-56 (host float, double 0.5, narrowed)
    volatile float a0 = g->d[0] * g->d[0], a1 = g->d[1] * g->d[1], a2 = g->d[2] * g->d[2];
    volatile float av = sqrtf((a0 + a1) + a2);
    volatile float b0 = s->segx * s->segx, b1 = s->segy * s->segy, b2 = s->segz * s->segz;
    volatile float bv = sqrtf((b0 + b1) + b2);
    volatile float sum = av + bv;
    G.sumRadii = (float)(0.5 * (double)sum);
    G.halfDiag2 = 0.25f * ((a0 + a1) + a2);
  }
  for (int a = 0; a < s->na; a++) {
    // MeasurementSetupTrafo2CartCoord.hpp:67-68: T(M_PI)*ida*(da)/T(180.0)
    volatile float t1 = (float)M_PI * (float)a;
    volatile float t2 = t1 * s->da;
    volatile float arg = t2 / 180.0f;
    G.sin_[a] = sinf(arg);
    G.cos_[a] = cosf(arg);
  }
  ctx->have_geo = true;
  return EPET_OK;
}

int epet_set_nrays(epet_ctx* ctx, int nrays) {
  EPET_CHECK(ctx, EPET_ERR_ARG, "null context");
  EPET_CHECK(nrays >= 1 && nrays <= EPET_MAX_RAYS, EPET_ERR_UNSUPPORTED,
             "rays per channel must be in [1, " + std::to_string(EPET_MAX_RAYS) + "]");
  ctx->nrays = nrays;
  return EPET_OK;
}

int epet_set_seed(epet_ctx* ctx, uint64_t seed) {
  EPET_CHECK(ctx, EPET_ERR_ARG, "null context");
  ctx->seed = seed;
  return EPET_OK;
}

int epet_malloc(epet_ctx* ctx, void** ptr, size_t bytes) {
  EPET_CHECK(ctx && ptr, EPET_ERR_ARG, "null argument");
  if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess) {
    (void)hipGetLastError();
    EPET_FAIL(EPET_ERR_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  }
  return EPET_OK;
}

int epet_free(epet_ctx* ctx, void* ptr) {
  EPET_CHECK(ctx, EPET_ERR_ARG, "null context");
  if (ptr) {
    EPET_HIP(hipStreamSynchronize(ctx->stream));
    EPET_HIP(hipFree(ptr));
  }
  return EPET_OK;
}

int epet_memcpy_h2d(epet_ctx* ctx, void* dst, const void* src, size_t bytes) {
  EPET_CHECK(ctx && (bytes == 0 || (dst && src)), EPET_ERR_ARG, "null argument");
  if (bytes) {
    EPET_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    EPET_HIP(hipStreamSynchronize(ctx->stream));
  }
  return EPET_OK;
}

int epet_memcpy_d2h(epet_ctx* ctx, void* dst, const void* src, size_t bytes) {
  EPET_CHECK(ctx && (bytes == 0 || (dst && src)), EPET_ERR_ARG, "null argument");
  if (bytes) {
    EPET_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    EPET_HIP(hipStreamSynchronize(ctx->stream));
  }
  return EPET_OK;
}

int epet_memcpy_d2d(epet_ctx* ctx, void* dst, const void* src, size_t bytes) {
  EPET_CHECK(ctx && (bytes == 0 || (dst && src)), EPET_ERR_ARG, "null argument");
  if (bytes) EPET_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return EPET_OK;
}

// wrappers.hpp:410-422
uint32_t epet_nInChunk(uint32_t idChunk, uint32_t nAll, uint32_t nFullChunk) {
  if (nFullChunk == 0) return 0;
  if (idChunk < (nAll + nFullChunk - 1) / nFullChunk) {
    if (idChunk < nAll / nFullChunk) return nFullChunk;
    return nAll % nFullChunk;
  }
  return 0;
}
// wrappers.hpp:427-433
uint32_t epet_chunkPtr(uint32_t chunkId, uint32_t nFullChunk) { return chunkId * nFullChunk; }
// wrappers.hpp:439-445
uint32_t epet_nChunks(unsigned long long maxNnz, unsigned long long maxNChunk) {
  if (maxNChunk == 0) return 0;
  return (uint32_t)((maxNnz + maxNChunk - 1) / maxNChunk);
}

}  // extern "C"
