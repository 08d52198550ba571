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

// ------------------------------------------------------------------ vector forward
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

template <typename VT> struct Vec4;
template <> struct Vec4<ValF32> {
  static __device__ __forceinline__ void get(const float* p, long long k, float v[4]) {
    const f32x4 t = __builtin_nontemporal_load((const f32x4*)(p + k));
    v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
  }
};
template <> struct Vec4<ValBF16> {
  static __device__ __forceinline__ void get(const uint16_t* p, long long k, float v[4]) {
    const u16x4 t = __builtin_nontemporal_load((const u16x4*)(p + k));
    v[0] = __uint_as_float((uint32_t)t.x << 16); v[1] = __uint_as_float((uint32_t)t.y << 16);
    v[2] = __uint_as_float((uint32_t)t.z << 16); v[3] = __uint_as_float((uint32_t)t.w << 16);
  }
};
template <> struct Vec4<ValF16> {
  static __device__ __forceinline__ void get(const uint16_t* p, long long k, float v[4]) {
    const u16x4 t = __builtin_nontemporal_load((const u16x4*)(p + k));
    v[0] = __half2float(__ushort_as_half(t.x)); v[1] = __half2float(__ushort_as_half(t.y));
    v[2] = __half2float(__ushort_as_half(t.z)); v[3] = __half2float(__ushort_as_half(t.w));
  }
};

// Per-lane partial of sum_k a_k x[col_k] over row entries [b, e) for a whole
// wavefront: scalar head up to a 4-entry boundary, 16-byte (col) and 16/8-byte
// (value) vector loads for the body, two quads in flight per lane, scalar tail.
template <typename VT>
__device__ __forceinline__ float row_dot_wave(long long b, long long e, int lane,
                                              const int* __restrict__ col,
                                              const typename VT::T* __restrict__ val,
                                              const float* __restrict__ x) {
  float acc0 = 0.0f, acc1 = 0.0f;
  long long kh = (b + 3) & ~3ll;
  if (kh > e) kh = e;
  if (b + lane < kh) acc0 += VT::get(val, b + lane) * x[ldcol(col, b + lane)];
  const long long nq = (e - kh) >> 2;
  long long q = lane;
  for (; q + WAVE < nq; q += 2 * WAVE) {
    const long long k0 = kh + 4 * q, k1 = kh + 4 * (q + WAVE);
    const i32x4 c0 = __builtin_nontemporal_load((const i32x4*)(col + k0));
    const i32x4 c1 = __builtin_nontemporal_load((const i32x4*)(col + k1));
    float v0[4], v1[4];
    Vec4<VT>::get(val, k0, v0);
    Vec4<VT>::get(val, k1, v1);
    acc0 += v0[0] * x[c0.x] + v0[1] * x[c0.y] + v0[2] * x[c0.z] + v0[3] * x[c0.w];
    acc1 += v1[0] * x[c1.x] + v1[1] * x[c1.y] + v1[2] * x[c1.z] + v1[3] * x[c1.w];
  }
  if (q < nq) {
    const long long k0 = kh + 4 * q;
    const i32x4 c0 = __builtin_nontemporal_load((const i32x4*)(col + k0));
    float v0[4];
    Vec4<VT>::get(val, k0, v0);
    acc0 += v0[0] * x[c0.x] + v0[1] * x[c0.y] + v0[2] * x[c0.z] + v0[3] * x[c0.w];
  }
  const long long kt = kh + (nq << 2);
  if (kt + lane < e) acc1 += VT::get(val, kt + lane) * x[ldcol(col, kt + lane)];
  return acc0 + acc1;
}

// Same sum with every load of a pass issued before any value is used: U
// quads per lane (256U entries per wavefront per pass) plus, on the first
// pass, the unaligned head and the tail.  Out-of-range slots load a clamped
// in-row address and are zeroed by a select at use, so no load sits behind a
// branch and a row of up to 256U entries costs one HBM and one L2 round trip.
// Requires b < e.
template <typename VT, int U>
__device__ __forceinline__ float row_dot_wave_u(long long b, long long e, int lane,
                                                const int* __restrict__ col,
                                                const typename VT::T* __restrict__ val,
                                                const float* __restrict__ x) {
  long long kh = (b + 3) & ~3ll;
  if (kh > e) kh = e;
  const long long nq = (e - kh) >> 2;
  const long long kt = kh + (nq << 2);
  const long long last = e - 1;
  // head and tail entries of this lane (one each at most)
  const bool hv = b + lane < kh, tv = kt + lane < e;
  const long long hk = hv ? b + lane : last, tk = tv ? kt + lane : last;
  const int hc = ldcol(col, hk), tc = ldcol(col, tk);
  const float ha = VT::get(val, hk), ta = VT::get(val, tk);
  float acc = 0.0f;
  for (long long q0 = 0; q0 < nq; q0 += (long long)U * WAVE) {
    i32x4 c[U];
    float v[U][4];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const long long q = q0 + lane + (long long)u * WAVE;
      ok[u] = q < nq;
      const long long k = kh + 4 * (ok[u] ? q : nq - 1);
      c[u] = __builtin_nontemporal_load((const i32x4*)(col + k));
      Vec4<VT>::get(val, k, v[u]);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const float s = v[u][0] * x[c[u].x] + v[u][1] * x[c[u].y] + v[u][2] * x[c[u].z] +
                      v[u][3] * x[c[u].w];
      acc += ok[u] ? s : 0.0f;
    }
  }
  const float hs = ha * x[hc], ts = ta * x[tc];
  acc += hv ? hs : 0.0f;
  acc += tv ? ts : 0.0f;
  return acc;
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
