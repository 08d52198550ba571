// sparse.hip -- COO sort, COO->CSR->ECSR, and library-matrix utilities.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <vector>

#include "epet_block.hpp"
#include "epet_internal.hpp"

namespace epet {

// ------------------------------------------------------------------ radix sort
// Stable LSD radix sort of (key, value) pairs with 8-bit digits.  A key packs
// (row, col) with flipped sign bits, so unsigned order = lexicographic signed
// (row, col) order, the order of cooIdLess (cooSort.hpp:26-44).  Digits on
// which all keys agree are skipped.
constexpr int RS_BLOCK = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_BLOCK * RS_ITEMS;
constexpr int RS_BINS = 256;

__global__ __launch_bounds__(256) void k_pack_keys(const int* __restrict__ row,
                                                   const int* __restrict__ col, long long n,
                                                   uint64_t* __restrict__ key) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  key[i] = ((uint64_t)((uint32_t)row[i] ^ 0x80000000u) << 32) |
           (uint64_t)((uint32_t)col[i] ^ 0x80000000u);
}

__global__ __launch_bounds__(256) void k_unpack_keys(const uint64_t* __restrict__ key,
                                                     long long n, int* __restrict__ row,
                                                     int* __restrict__ col) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = key[i];
  row[i] = (int)((uint32_t)(k >> 32) ^ 0x80000000u);
  col[i] = (int)((uint32_t)k ^ 0x80000000u);
}

// bitwise AND and OR over all keys: digits where AND == OR are constant
__global__ __launch_bounds__(256) void k_key_andor(const uint64_t* __restrict__ key, long long n,
                                                   unsigned long long* __restrict__ andor) {
  uint64_t a = ~0ull, o = 0ull;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    a &= key[i];
    o |= key[i];
  }
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) {
    a &= __shfl_xor(a, off, WAVE);
    o |= __shfl_xor(o, off, WAVE);
  }
  if (lane_id() == 0) {
    atomicAnd(&andor[0], (unsigned long long)a);
    atomicOr(&andor[1], (unsigned long long)o);
  }
}

__global__ void k_init_andor(unsigned long long* andor) {
  if (threadIdx.x == 0) { andor[0] = ~0ull; andor[1] = 0ull; }
}

__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const uint64_t* __restrict__ key,
                                                      long long n, int shift, int ntiles,
                                                      int* __restrict__ hist) {
  __shared__ int h[RS_BINS];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * RS_TILE;
#pragma unroll
  for (int j = 0; j < RS_ITEMS; j++) {
    const long long i = base + j * RS_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&h[(key[i] >> shift) & 0xFF], 1);
  }
  __syncthreads();
  hist[(long long)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const uint64_t* __restrict__ kin,
                                                         const float* __restrict__ vin,
                                                         long long n, int shift, int ntiles,
                                                         const int64_t* __restrict__ offs,
                                                         uint64_t* __restrict__ kout,
                                                         float* __restrict__ vout) {
  __shared__ int64_t s_base[RS_BINS];
  __shared__ int s_run[RS_BINS];
  __shared__ int s_wcnt[RS_BLOCK / WAVE][RS_BINS];
  const int tid = threadIdx.x;
  s_base[tid] = offs[(long long)tid * ntiles + blockIdx.x];
  s_run[tid] = 0;
  const int w = wave_id(), lane = lane_id();
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  const long long base = (long long)blockIdx.x * RS_TILE;
  for (int j = 0; j < RS_ITEMS; j++) {
#pragma unroll
    for (int q = 0; q < RS_BLOCK / WAVE; q++) s_wcnt[q][tid] = 0;
    __syncthreads();
    const long long i = base + j * RS_BLOCK + tid;
    const bool valid = i < n;
    uint64_t k = 0;
    float v = 0.0f;
    if (valid) { k = kin[i]; v = vin[i]; }
    const int d = (int)((k >> shift) & 0xFF);
    // lanes of this wavefront holding the same digit
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const bool bit = (d >> b) & 1;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const int rank = __popcll(peers & lt);
    if (valid && rank == 0) s_wcnt[w][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int before = 0;
      for (int q = 0; q < w; q++) before += s_wcnt[q][d];
      const int64_t pos = s_base[d] + s_run[d] + before + rank;
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int q = 0; q < RS_BLOCK / WAVE; q++) add += s_wcnt[q][tid];
    s_run[tid] += add;
    __syncthreads();
  }
}

// ------------------------------------------------------------------ CSR helpers
__global__ __launch_bounds__(256) void k_fill_i32(int* __restrict__ p, long long n, int v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// row ids outside [0, m) are not counted (and flagged)
__global__ __launch_bounds__(256) void k_row_hist(const int* __restrict__ rowid, long long nnz,
                                                  int m, int* __restrict__ cnt,
                                                  int* __restrict__ bad) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nnz;
       i += (long long)gridDim.x * blockDim.x) {
    const int r = rowid[i];
    if (r >= 0 && r < m) atomicAdd(&cnt[r], 1);
    else atomicOr(bad, 1);
  }
}

__global__ __launch_bounds__(256) void k_csr2ecsr(int* __restrict__ ecsr,
                                                  const int* __restrict__ yRowId, int nnzr,
                                                  const int* __restrict__ csr, int m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nnzr) {
    const int y = yRowId[i];
    ecsr[i] = (y >= 0 && y <= m) ? csr[y] : -1;   // -1 marks an out-of-range row id
  }
  if (i == nnzr) ecsr[nnzr] = csr[m];
}

__global__ __launch_bounds__(256) void k_sel_counts(const int64_t* __restrict__ rowptr,
                                                    const int* __restrict__ rows, int nrows,
                                                    int m, int* __restrict__ cnt,
                                                    int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const int r = rows[i];
  if (r < 0 || r >= m) { atomicOr(bad, 1); cnt[i] = 0; return; }
  cnt[i] = (int)(rowptr[r + 1] - rowptr[r]);
}

template <typename V>
__global__ __launch_bounds__(256) void k_sel_copy(const int64_t* __restrict__ rowptr,
                                                  const int* __restrict__ col,
                                                  const V* __restrict__ val,
                                                  const float* __restrict__ scale,
                                                  const int* __restrict__ rows, int m,
                                                  const int64_t* __restrict__ off,
                                                  int64_t* __restrict__ orow,
                                                  int* __restrict__ ocol, V* __restrict__ oval,
                                                  float* __restrict__ oscale) {
  const int i = blockIdx.x;
  const int r = rows[i];
  const int64_t o = off[i];
  if (threadIdx.x == 0) {
    orow[i] = o;
    if (oscale) oscale[i] = (r >= 0 && r < m) ? scale[r] : 0.0f;
  }
  if (r < 0 || r >= m) return;
  const int64_t b = rowptr[r], e = rowptr[r + 1];
  for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
    ocol[o + (k - b)] = col[k];
    oval[o + (k - b)] = val[k];
  }
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// per-row scale = max |a|; stored value = a / scale
__global__ __launch_bounds__(256) void k_convert_rows(const int64_t* __restrict__ rowptr,
                                                      const float* __restrict__ val, int vt,
                                                      uint16_t* __restrict__ out,
                                                      float* __restrict__ scale) {
  __shared__ float s_w[256 / WAVE];
  const int r = blockIdx.x;
  const int64_t b = rowptr[r], e = rowptr[r + 1];
  float mx = 0.0f;
  for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) mx = fmaxf(mx, fabsf(val[k]));
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, WAVE));
  if (lane_id() == 0) s_w[wave_id()] = mx;
  __syncthreads();
  mx = fmaxf(fmaxf(s_w[0], s_w[1]), fmaxf(s_w[2], s_w[3]));
  const float sc = mx > 0.0f ? mx : 1.0f;
  if (threadIdx.x == 0) scale[r] = sc;
  for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
    const float q = val[k] / sc;
    out[k] = (vt == EPET_VAL_BF16) ? f32_to_bf16_rne(q) : __half_as_ushort(__float2half_rn(q));
  }
}

template <typename V>
__device__ __forceinline__ float load_val(const V* p, long long k);
template <>
__device__ __forceinline__ float load_val<float>(const float* p, long long k) { return p[k]; }

__global__ __launch_bounds__(256) void k_colsum_f64(const int64_t* __restrict__ rowptr,
                                                    const int* __restrict__ col,
                                                    const void* __restrict__ val, int vt,
                                                    const float* __restrict__ scale, int m,
                                                    double* __restrict__ s) {
  const int r = blockIdx.x;
  if (r >= m) return;
  const int64_t b = rowptr[r], e = rowptr[r + 1];
  const float sc = scale ? scale[r] : 1.0f;
  for (int64_t k = b + threadIdx.x; k < e; k += blockDim.x) {
    float a;
    if (vt == EPET_VAL_F32) a = ((const float*)val)[k];
    else if (vt == EPET_VAL_BF16)
      a = __uint_as_float((uint32_t)((const uint16_t*)val)[k] << 16) * sc;
    else a = __half2float(__ushort_as_half(((const uint16_t*)val)[k])) * sc;
    unsafeAtomicAdd(&s[col[k]], (double)a);
  }
}

__global__ __launch_bounds__(256) void k_fill_f64(double* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_set_last_i64(int64_t* p, const int64_t* v) {
  if (threadIdx.x == 0) *p = *v;
}

__global__ __launch_bounds__(256) void k_gather_i32(const int* __restrict__ src,
                                                    const int* __restrict__ idx, int n,
                                                    int* __restrict__ dst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// Structure check of caller arrays: bit 1 = invalid (row pointers not
// monotone / not spanning [0, nnz], column outside [0, n)); bit 2 = columns not
// strictly ascending within a row.
__global__ __launch_bounds__(256) void k_check_csr(int m, int n, long long nnz,
                                                   const int64_t* __restrict__ rowptr,
                                                   const int* __restrict__ col,
                                                   int* __restrict__ flags) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const int64_t b = rowptr[r], e = rowptr[r + 1];
  int f = 0;
  if (r == 0 && b != 0) f |= 1;
  if (r == m - 1 && e != nnz) f |= 1;
  if (e < b || b < 0 || e > nnz) f |= 1;
  else {
    int prev = -1;
    for (int64_t k = b; k < e; k++) {
      const int c = col[k];
      if (c < 0 || c >= n) { f |= 1; break; }
      if (c <= prev) f |= 2;
      prev = c;
    }
  }
  if (f) atomicOr(flags, f);
}

}  // namespace epet

using namespace epet;

namespace epet {
// ------------------------------------------------------------------ transpose
// Deterministic transpose without atomics: workgroup b walks its block of rows
// in order, one row at a time (columns are distinct within a row, a barrier
// separates rows), counting entries per (column, block) and later placing them
// at column offset + block offset + running count.  Entries of a column come
// out in ascending row order.
constexpr int TR_BLOCK = 256;

__global__ __launch_bounds__(TR_BLOCK) void k_tr_count(int m, int rows_per_block, int B,
                                                       const int64_t* __restrict__ rowptr,
                                                       const int* __restrict__ col,
                                                       int* __restrict__ cnt) {
  const int blk = blockIdx.x;
  const long long r0 = (long long)blk * rows_per_block;
  const long long r1 = min((long long)m, r0 + rows_per_block);
  for (long long r = r0; r < r1; r++) {
    const int64_t b = rowptr[r], e = rowptr[r + 1];
    for (int64_t k = b + threadIdx.x; k < e; k += TR_BLOCK) {
      int* p = &cnt[(size_t)col[k] * B + blk];
      *p = *p + 1;
    }
    __syncthreads();
  }
}

template <typename V>
__global__ __launch_bounds__(TR_BLOCK) void k_tr_fill(int m, int rows_per_block, int B,
                                                      const int64_t* __restrict__ rowptr,
                                                      const int* __restrict__ col,
                                                      const V* __restrict__ val,
                                                      int64_t* __restrict__ cursor,
                                                      int* __restrict__ orow,
                                                      V* __restrict__ oval) {
  const int blk = blockIdx.x;
  const long long r0 = (long long)blk * rows_per_block;
  const long long r1 = min((long long)m, r0 + rows_per_block);
  for (long long r = r0; r < r1; r++) {
    const int64_t b = rowptr[r], e = rowptr[r + 1];
    for (int64_t k = b + threadIdx.x; k < e; k += TR_BLOCK) {
      int64_t* p = &cursor[(size_t)col[k] * B + blk];
      const int64_t pos = *p;
      *p = pos + 1;
      orow[pos] = (int)r;
      oval[pos] = val[k];
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_tr_rowptr(int n, int B, long long nnz,
                                                   const int64_t* __restrict__ off,
                                                   int64_t* __restrict__ tptr) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) tptr[j] = off[(size_t)j * B];
  if (j == n) tptr[n] = nnz;
}
}  // namespace epet

epet_matrix::~epet_matrix() {
  if (bp_toff) (void)hipFree(bp_toff);
  if (bp_partial) (void)hipFree(bp_partial);
  if (xb_toff) (void)hipFree(xb_toff);
  if (xb_partial) (void)hipFree(xb_partial);
  if (lds_off) (void)hipFree(lds_off);
  if (lds_val) (void)hipFree(lds_val);
  if (lds_col16) (void)hipFree(lds_col16);
  if (lds_partial) (void)hipFree(lds_partial);
  epet::blk_release(this);
  if (tr) epet_matrix_destroy(tr);
}

namespace epet {
void xb_prepare(epet_ctx* ctx, epet_matrix* At);
int matrix_alloc(epet_matrix** out, int m, int n, long long nnz, int value_type) {
  epet_matrix* A = new epet_matrix();
  A->m = m;
  A->n = n;
  A->nnz = nnz;
  A->value_type = value_type;
  const size_t vb = value_type == EPET_VAL_F32 ? 4 : 2;
  bool ok = hipMalloc((void**)&A->rowptr, sizeof(int64_t) * ((size_t)m + 1)) == hipSuccess &&
            hipMalloc((void**)&A->col, sizeof(int) * (size_t)std::max(1ll, nnz)) == hipSuccess &&
            hipMalloc(&A->val, vb * (size_t)std::max(1ll, nnz)) == hipSuccess;
  if (ok && value_type != EPET_VAL_F32)
    ok = hipMalloc((void**)&A->rowscale, sizeof(float) * (size_t)std::max(1, m)) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    if (A->rowptr) (void)hipFree(A->rowptr);
    if (A->col) (void)hipFree(A->col);
    if (A->val) (void)hipFree(A->val);
    delete A;
    EPET_FAIL(EPET_ERR_ALLOC, "matrix allocation (" + std::to_string(nnz) + " entries) failed");
  }
  (void)hipGetDevice(&A->device);
  *out = A;
  return EPET_OK;
}
}  // namespace epet

extern "C" {

int epet_cooSort(epet_ctx* ctx, float* val, int* row, int* col, long long N) {
  EPET_CHECK(ctx && (N == 0 || (val && row && col)), EPET_ERR_ARG, "null argument");
  EPET_CHECK(N >= 0, EPET_ERR_ARG, "negative length");
  if (N <= 1) return EPET_OK;
  const long long ntiles = (N + RS_TILE - 1) / RS_TILE;
  EPET_CHECK(ntiles * RS_BINS < (1ll << 31), EPET_ERR_UNSUPPORTED, "cooSort input too long");
  uint64_t *k0 = nullptr, *k1 = nullptr;
  float* v1 = nullptr;
  int* hist = nullptr;
  int64_t* offs = nullptr;
  unsigned long long* andor = nullptr;
  hipStream_t st = ctx->stream;
  StreamTemps tmp(st);
  EPET_HIP(tmp.alloc(&k0, sizeof(uint64_t) * N));
  EPET_HIP(tmp.alloc(&k1, sizeof(uint64_t) * N));
  EPET_HIP(tmp.alloc(&v1, sizeof(float) * N));
  EPET_HIP(tmp.alloc(&hist, sizeof(int) * ntiles * RS_BINS));
  EPET_HIP(tmp.alloc(&offs, sizeof(int64_t) * (ntiles * RS_BINS + 1)));
  EPET_HIP(tmp.alloc(&andor, sizeof(unsigned long long) * 2));
  const unsigned g = grid_for(N, 256, 1ll << 31);
  k_pack_keys<<<g, 256, 0, st>>>(row, col, N, k0);
  k_init_andor<<<1, 64, 0, st>>>(andor);
  k_key_andor<<<grid_for(N, 256, 4096), 256, 0, st>>>(k0, N, andor);
  EPET_HIP(hipGetLastError());
  unsigned long long ao[2];
  EPET_TRY(read_pinned(ctx, ao, andor, sizeof(ao)));
  const uint64_t vary = ao[0] ^ ao[1];
  uint64_t* kin = k0;
  uint64_t* kout = k1;
  float* vin = val;
  float* vout = v1;
  int rc = EPET_OK;
  for (int shift = 0; shift < 64 && rc == EPET_OK; shift += 8) {
    if (((vary >> shift) & 0xFF) == 0) continue;
    k_rs_hist<<<(unsigned)ntiles, RS_BLOCK, 0, st>>>(kin, N, shift, (int)ntiles, hist);
    EPET_HIP(hipGetLastError());
    rc = exclusive_scan_i32_to_i64(ctx, hist, offs, ntiles * RS_BINS, nullptr);
    k_rs_scatter<<<(unsigned)ntiles, RS_BLOCK, 0, st>>>(kin, vin, N, shift, (int)ntiles, offs,
                                                        kout, vout);
    EPET_HIP(hipGetLastError());
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  if (vin != val) EPET_HIP(hipMemcpyAsync(val, vin, sizeof(float) * N, hipMemcpyDeviceToDevice, st));
  k_unpack_keys<<<g, 256, 0, st>>>(kin, N, row, col);
  EPET_HIP(hipGetLastError());
  return rc;
}

int epet_convertCoo2Csr(epet_ctx* ctx, int* csrRowPtr, const int* cooRowId, long long nnz, int m) {
  EPET_CHECK(ctx && csrRowPtr && (nnz == 0 || cooRowId), EPET_ERR_ARG, "null argument");
  EPET_CHECK(m >= 0 && nnz >= 0 && nnz < (1ll << 31), EPET_ERR_ARG, "bad sizes");
  hipStream_t st = ctx->stream;
  StreamTemps tmp(st);
  int* bad = nullptr;
  EPET_HIP(tmp.alloc(&bad, sizeof(int)));
  k_fill_i32<<<1, 64, 0, st>>>(bad, 1, 0);
  k_fill_i32<<<grid_for(m + 1, 256, 1ll << 31), 256, 0, st>>>(csrRowPtr, m + 1, 0);
  if (nnz > 0)
    k_row_hist<<<grid_for(nnz, 256, 65536), 256, 0, st>>>(cooRowId, nnz, m, csrRowPtr, bad);
  EPET_HIP(hipGetLastError());
  EPET_TRY(exclusive_scan_i32(ctx, csrRowPtr, csrRowPtr, (long long)m + 1, nullptr));
  int hbad = 0;
  EPET_TRY(read_pinned(ctx, &hbad, bad, sizeof(int)));
  EPET_CHECK(!hbad, EPET_ERR_ARG, "COO row id outside [0, m)");
  return EPET_OK;
}

int epet_convertCsr2Ecsr(epet_ctx* ctx, int* ecsr, const int* yRowId, int nnzr, const int* csr,
                         int m) {
  EPET_CHECK(ctx && ecsr && csr && (nnzr == 0 || yRowId), EPET_ERR_ARG, "null argument");
  EPET_CHECK(nnzr >= 0 && m >= 0, EPET_ERR_ARG, "bad sizes");
  k_csr2ecsr<<<grid_for((long long)nnzr + 1, 256, 1ll << 31), 256, 0, ctx->stream>>>(ecsr, yRowId,
                                                                                  nnzr, csr, m);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

int epet_matrix_info(const epet_matrix* A, int* m, int* n, long long* nnz, int* value_type) {
  EPET_CHECK(A, EPET_ERR_ARG, "null matrix");
  if (m) *m = A->m;
  if (n) *n = A->n;
  if (nnz) *nnz = A->nnz;
  if (value_type) *value_type = A->value_type;
  return EPET_OK;
}

int epet_matrix_arrays(const epet_matrix* A, int64_t** rowptr, int** col, void** val,
                       float** rowscale) {
  EPET_CHECK(A, EPET_ERR_ARG, "null matrix");
  if (rowptr) *rowptr = A->rowptr;
  if (col) *col = A->col;
  if (val) *val = A->val;
  if (rowscale) *rowscale = A->rowscale;
  return EPET_OK;
}

int epet_matrix_destroy(epet_matrix* A) {
  if (!A) return EPET_OK;
  (void)hipDeviceSynchronize();
  if (A->owns) {
    if (A->rowptr) (void)hipFree(A->rowptr);
    if (A->col) (void)hipFree(A->col);
    if (A->val) (void)hipFree(A->val);
    if (A->rowscale) (void)hipFree(A->rowscale);
    if (A->row_ids) (void)hipFree(A->row_ids);
  }
  delete A;
  return EPET_OK;
}

int epet_matrix_create(epet_ctx* ctx, int m, int n, long long nnz, const int64_t* rowptr,
                       const int* col, const float* val, epet_matrix** out) {
  EPET_CHECK(ctx && out && rowptr && (nnz == 0 || (col && val)), EPET_ERR_ARG, "null argument");
  EPET_CHECK(m >= 0 && n >= 0 && nnz >= 0, EPET_ERR_ARG, "bad sizes");
  hipStream_t st = ctx->stream;
  int hflags = 0;
  if (m > 0) {
    StreamTemps tmp(st);
    int* flags = nullptr;
    EPET_HIP(tmp.alloc(&flags, sizeof(int)));
    k_fill_i32<<<1, 64, 0, st>>>(flags, 1, 0);
    k_check_csr<<<grid_for(m, 256, 1ll << 31), 256, 0, st>>>(m, n, nnz, rowptr, col, flags);
    EPET_HIP(hipGetLastError());
    EPET_TRY(read_pinned(ctx, &hflags, flags, sizeof(int)));
    EPET_CHECK(!(hflags & 1), EPET_ERR_ARG,
               "invalid CSR arrays (row pointers or column ids out of range)");
  }
  epet_matrix* A = nullptr;
  EPET_TRY(matrix_alloc(&A, m, n, nnz, EPET_VAL_F32));
  MatrixGuard keep(A);
  // the tiled back projection needs ascending columns within rows
  if (hflags & 2) { A->bp_failed = true; A->cols_strict = false; }
  EPET_HIP(hipMemcpyAsync(A->rowptr, rowptr, sizeof(int64_t) * ((size_t)m + 1),
                          hipMemcpyDeviceToDevice, st));
  if (nnz) {
    EPET_HIP(hipMemcpyAsync(A->col, col, sizeof(int) * nnz, hipMemcpyDeviceToDevice, st));
    EPET_HIP(hipMemcpyAsync(A->val, val, sizeof(float) * nnz, hipMemcpyDeviceToDevice, st));
  }
  EPET_HIP(hipStreamSynchronize(st));
  *out = keep.release();
  return EPET_OK;
}

int epet_matrix_select_rows(epet_ctx* ctx, const epet_matrix* A, const int* rows, int nrows,
                            epet_matrix** out) {
  EPET_CHECK(ctx && A && out && (nrows == 0 || rows), EPET_ERR_ARG, "null argument");
  EPET_CHECK(nrows >= 0, EPET_ERR_ARG, "negative row count");
  hipStream_t st = ctx->stream;
  StreamTemps tmp(st);
  int* cnt = nullptr;
  int64_t* off = nullptr;
  int* bad = nullptr;
  EPET_HIP(tmp.alloc(&cnt, sizeof(int) * std::max(1, nrows)));
  EPET_HIP(tmp.alloc(&off, sizeof(int64_t) * ((size_t)nrows + 1)));
  EPET_HIP(tmp.alloc(&bad, sizeof(int)));
  k_fill_i32<<<1, 64, 0, st>>>(bad, 1, 0);
  if (nrows)
    k_sel_counts<<<grid_for(nrows, 256, 1ll << 31), 256, 0, st>>>(A->rowptr, rows, nrows, A->m,
                                                                  cnt, bad);
  EPET_HIP(hipGetLastError());
  int hbad = 0;
  EPET_TRY(read_pinned(ctx, &hbad, bad, sizeof(int)));
  EPET_CHECK(!hbad, EPET_ERR_ARG, "row index out of range");
  long long nnz = 0;
  EPET_TRY(exclusive_scan_i32_to_i64(ctx, cnt, off, nrows, &nnz));
  epet_matrix* B = nullptr;
  EPET_TRY(matrix_alloc(&B, nrows, A->n, nnz, A->value_type));
  MatrixGuard keep(B);
  if (nrows) {
    if (A->value_type == EPET_VAL_F32)
      k_sel_copy<float><<<nrows, 256, 0, st>>>(A->rowptr, A->col, (const float*)A->val,
                                               A->rowscale, rows, A->m, off, B->rowptr, B->col,
                                               (float*)B->val, B->rowscale);
    else
      k_sel_copy<uint16_t><<<nrows, 256, 0, st>>>(A->rowptr, A->col, (const uint16_t*)A->val,
                                                  A->rowscale, rows, A->m, off, B->rowptr,
                                                  B->col, (uint16_t*)B->val, B->rowscale);
    EPET_HIP(hipGetLastError());
  }
  if (A->row_ids) {
    if (hipMalloc((void**)&B->row_ids, sizeof(int) * (size_t)std::max(1, nrows)) != hipSuccess) {
      (void)hipGetLastError();
      B->row_ids = nullptr;
    } else if (nrows) {
      k_gather_i32<<<grid_for(nrows, 256, 1ll << 31), 256, 0, st>>>(A->row_ids, rows, nrows,
                                                                    B->row_ids);
      EPET_HIP(hipGetLastError());
    }
    if (B->row_ids) B->mg = A->mg;
  }
  // last row pointer = nnz
  {
    int64_t h = nnz;
    EPET_HIP(hipMemcpyAsync(B->rowptr + nrows, &h, sizeof(int64_t), hipMemcpyHostToDevice, st));
    EPET_HIP(hipStreamSynchronize(st));
  }
  *out = keep.release();
  return EPET_OK;
}

int epet_matrix_convert(epet_ctx* ctx, const epet_matrix* A, int vt, epet_matrix** out) {
  EPET_CHECK(ctx && A && out, EPET_ERR_ARG, "null argument");
  EPET_CHECK(A->value_type == EPET_VAL_F32, EPET_ERR_STATE, "conversion source must be FP32");
  EPET_CHECK(vt == EPET_VAL_BF16 || vt == EPET_VAL_F16 || vt == EPET_VAL_F32, EPET_ERR_ARG,
             "unknown value type");
  hipStream_t st = ctx->stream;
  epet_matrix* B = nullptr;
  EPET_TRY(matrix_alloc(&B, A->m, A->n, A->nnz, vt));
  MatrixGuard keep(B);
  EPET_HIP(hipMemcpyAsync(B->rowptr, A->rowptr, sizeof(int64_t) * ((size_t)A->m + 1),
                          hipMemcpyDeviceToDevice, st));
  if (A->nnz) EPET_HIP(hipMemcpyAsync(B->col, A->col, sizeof(int) * A->nnz, hipMemcpyDeviceToDevice, st));
  if (vt == EPET_VAL_F32) {
    if (A->nnz)
      EPET_HIP(hipMemcpyAsync(B->val, A->val, sizeof(float) * A->nnz, hipMemcpyDeviceToDevice, st));
  } else if (A->m) {
    k_convert_rows<<<A->m, 256, 0, st>>>(A->rowptr, (const float*)A->val, vt,
                                         (uint16_t*)B->val, B->rowscale);
    EPET_HIP(hipGetLastError());
  }
  if (A->row_ids && hipMalloc((void**)&B->row_ids, sizeof(int) * (size_t)std::max(1, A->m)) ==
                        hipSuccess) {
    if (A->m)
      EPET_HIP(hipMemcpyAsync(B->row_ids, A->row_ids, sizeof(int) * A->m, hipMemcpyDeviceToDevice,
                              st));
    B->mg = A->mg;
  } else {
    (void)hipGetLastError();
  }
  EPET_HIP(hipStreamSynchronize(st));
  *out = keep.release();
  return EPET_OK;
}

int epet_matrix_colsum_f64(epet_ctx* ctx, const epet_matrix* A, double* s) {
  EPET_CHECK(ctx && A && s, EPET_ERR_ARG, "null argument");
  hipStream_t st = ctx->stream;
  k_fill_f64<<<grid_for(A->n, 256, 1ll << 31), 256, 0, st>>>(s, A->n);
  if (A->m)
    k_colsum_f64<<<A->m, 256, 0, st>>>(A->rowptr, A->col, A->val, A->value_type, A->rowscale,
                                       A->m, s);
  EPET_HIP(hipGetLastError());
  return EPET_OK;
}

int epet_matrix_transpose(epet_ctx* ctx, epet_matrix* A) {
  EPET_CHECK(ctx && A, EPET_ERR_ARG, "null argument");
  if (A->tr) return EPET_OK;
  EPET_CHECK(A->cols_strict, EPET_ERR_STATE,
             "transpose needs strictly ascending columns within every row");
  hipStream_t st = ctx->stream;
  const int rows_per_block = 4096;
  const int B = std::max(1, (A->m + rows_per_block - 1) / rows_per_block);
  const long long ncnt = (long long)A->n * B;
  StreamTemps tmp(st);
  int* cnt = nullptr;
  int64_t* off = nullptr;
  EPET_HIP(tmp.alloc(&cnt, sizeof(int) * std::max(1ll, ncnt)));
  EPET_HIP(tmp.alloc(&off, sizeof(int64_t) * (ncnt + 1)));
  k_fill_i32<<<grid_for(ncnt, 256, 1ll << 31), 256, 0, st>>>(cnt, ncnt, 0);
  if (A->m) k_tr_count<<<B, TR_BLOCK, 0, st>>>(A->m, rows_per_block, B, A->rowptr, A->col, cnt);
  EPET_HIP(hipGetLastError());
  long long tot = 0;
  EPET_TRY(exclusive_scan_i32_to_i64(ctx, cnt, off, ncnt, &tot));
  EPET_CHECK(tot == A->nnz, EPET_ERR_STATE, "transpose count mismatch");
  epet_matrix* T = nullptr;
  EPET_TRY(matrix_alloc(&T, A->n, A->m, A->nnz, A->value_type == EPET_VAL_F32 ? EPET_VAL_F32
                                                                              : A->value_type));
  MatrixGuard keep(T);
  if (T->rowscale) { (void)hipFree(T->rowscale); T->rowscale = nullptr; }
  k_tr_rowptr<<<grid_for((long long)A->n + 1, 256, 1ll << 31), 256, 0, st>>>(A->n, B, A->nnz, off,
                                                                           T->rowptr);
  if (A->m) {
    if (A->value_type == EPET_VAL_F32)
      k_tr_fill<float><<<B, TR_BLOCK, 0, st>>>(A->m, rows_per_block, B, A->rowptr, A->col,
                                               (const float*)A->val, off, T->col, (float*)T->val);
    else
      k_tr_fill<uint16_t><<<B, TR_BLOCK, 0, st>>>(A->m, rows_per_block, B, A->rowptr, A->col,
                                                  (const uint16_t*)A->val, off, T->col,
                                                  (uint16_t*)T->val);
  }
  EPET_HIP(hipGetLastError());
  EPET_HIP(hipStreamSynchronize(st));
  T->bp_failed = true;
  xb_prepare(ctx, T);   // optional: without a plan the plain row kernel runs
  A->tr = keep.release();
  return EPET_OK;
}

}  // extern "C"
