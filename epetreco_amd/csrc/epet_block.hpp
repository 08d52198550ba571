// epet_block.hpp -- wavefront/workgroup primitives for 64-wide wavefronts.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace epet {

constexpr int WAVE = 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / WAVE; }

// Rank of this lane among the lanes of its wavefront whose flag is set, and
// the wavefront's count (ballot + popcount: order-preserving compaction).
__device__ __forceinline__ int wave_rank(bool flag, int* count) {
  const uint64_t m = __ballot(flag);
  const int lane = lane_id();
  *count = __popcll(m);
  return __popcll(m & ((lane == 0) ? 0ull : (~0ull >> (64 - lane))));
}

// Workgroup-wide exclusive scan of a 0/1 flag in thread order.  s_wave must
// hold BLOCK/64 ints of LDS.  Returns this thread's position; *total = count.
// Contains two barriers: every thread of the block must call it.
template <int BLOCK>
__device__ __forceinline__ int block_rank(bool flag, int* s_wave, int* total) {
  int wcount;
  const int r = wave_rank(flag, &wcount);
  if (lane_id() == 0) s_wave[wave_id()] = wcount;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < BLOCK / WAVE; w++) {
    const int c = s_wave[w];
    base += (w < wave_id()) ? c : 0;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return base + r;
}

// Workgroup-wide exclusive scan of an int value in thread order.
template <int BLOCK>
__device__ __forceinline__ int block_excl_sum(int v, int* s_wave, int* total) {
  // inclusive scan within the wavefront
  int x = v;
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < WAVE; off <<= 1) {
    const int y = __shfl_up(x, off, WAVE);
    if (lane >= off) x += y;
  }
  if (lane == WAVE - 1) s_wave[wave_id()] = x;
  __syncthreads();
  int base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < BLOCK / WAVE; w++) {
    const int c = s_wave[w];
    base += (w < wave_id()) ? c : 0;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
  return v;
}

}  // namespace epet
