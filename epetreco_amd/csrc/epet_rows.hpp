// epet_rows.hpp -- value decoding and the wavefront row dot product shared by
// the projection kernels.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include "epet_block.hpp"

namespace epet {

// stored value -> float (reduced-precision values are multiplied by the row
// scale by the caller)
// get: non-temporal (streaming) load; get_cached: ordinary load
struct ValF32 {
  using T = float;
  static __device__ __forceinline__ float get(const T* p, long long k) {
    return __builtin_nontemporal_load(p + k);
  }
  static __device__ __forceinline__ float get_cached(const T* p, long long k) { return p[k]; }
};
struct ValBF16 {
  using T = uint16_t;
  static __device__ __forceinline__ float get(const T* p, long long k) {
    return __uint_as_float((uint32_t)__builtin_nontemporal_load(p + k) << 16);
  }
  static __device__ __forceinline__ float get_cached(const T* p, long long k) {
    return __uint_as_float((uint32_t)p[k] << 16);
  }
};
struct ValF16 {
  using T = uint16_t;
  static __device__ __forceinline__ float get(const T* p, long long k) {
    return __half2float(__ushort_as_half(__builtin_nontemporal_load(p + k)));
  }
  static __device__ __forceinline__ float get_cached(const T* p, long long k) {
    return __half2float(__ushort_as_half(p[k]));
  }
};

// Row bounds prefetch: lanes 0 and 1 load rowptr[r] and rowptr[r + 1] in one
// vector load (vector loads complete in order, so waiting for the row in
// flight never waits for this younger request); row_bounds() reads them out.
__device__ __forceinline__ long long row_bounds_load(const int64_t* __restrict__ rowptr,
                                                     long long r, int lane) {
  return rowptr[r + (lane & 1)];
}
__device__ __forceinline__ long long lane_value_i64(long long v, int l) {
  const int lo = __builtin_amdgcn_readlane((int)(unsigned long long)v, l);
  const int hi = __builtin_amdgcn_readlane((int)((unsigned long long)v >> 32), l);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

template <typename IDX>
__device__ __forceinline__ int ldcol(const int* p, IDX k) {
  return __builtin_nontemporal_load(p + k);
}

// Same sum with the lane-strided layout: lane l takes entries base + l + 64 j,
// j < J, so each gather instruction reads x at 64 consecutive entries of the
// row (neighbouring voxels, few distinct cache lines); all 2J loads of a pass
// are issued before the first gather.  Slots past the row end load the last
// entry and are zeroed by a select.  Requires b < e.
template <typename VT, int J>
__device__ __forceinline__ float row_dot_wave_s(long long b, long long e, int lane,
                                                const int* __restrict__ col,
                                                const typename VT::T* __restrict__ val,
                                                const float* __restrict__ x) {
  float acc = 0.0f;
  const long long last = e - 1;
  for (long long base = b; base < e; base += (long long)J * WAVE) {
    int c[J];
    float v[J];
    bool ok[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
      const long long k = base + lane + (long long)j * WAVE;
      ok[j] = k < e;
      const long long kc = ok[j] ? k : last;
      c[j] = ldcol(col, kc);
      v[j] = VT::get(val, kc);
    }
    float g[J];
#pragma unroll
    for (int j = 0; j < J; j++) g[j] = x[c[j]];
#pragma unroll
    for (int j = 0; j < J; j++) acc += ok[j] ? v[j] * g[j] : 0.0f;
  }
  return acc;
}

}  // namespace epet
