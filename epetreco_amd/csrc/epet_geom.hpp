// epet_geom.hpp -- device geometry of the system-matrix generator.
//
// Every function evaluates the reference's single-precision arithmetic in the
// reference's order, so device results are bit-identical with the host (the
// translation unit is compiled with -ffp-contract=off and correctly rounded
// division and square root).  The only host-side quantity is the per-angle
// sin/cos table, computed by epet_set_geometry with the reference's argument
// expression.
#pragma once

#include "epet_internal.hpp"

#pragma clang fp contract(off)

namespace epet {

// Channel id -> (ida, id0z, id0y, id1z, id1y) (MeasurementSetupLinIndex.hpp:58-165).
struct ChannelIds { int ida, id0z, id0y, id1z, id1y; };

__device__ __forceinline__ ChannelIds channel_ids(const Geo& g, int cnl) {
  ChannelIds c;
  int t = cnl % g.P4;
  c.ida = cnl / g.P4;
  c.id0z = t / g.P3;
  t %= g.P3;
  c.id0y = t / g.P2;
  t %= g.P2;
  c.id1z = t / g.P1;
  c.id1y = t % g.P1;
  return c;
}

// Segment box coordinates -> world coordinates for detector det
// (MeasurementSetupTrafo2CartCoord.hpp:56-104 and :225-273).
__device__ __forceinline__ void trafo(const Geo& g, int det, float b0, float b1, float b2,
                                      int idz, int idy, int ida, float out[3]) {
  const float e0 = g.segx, e1 = g.segy, e2 = g.segz;
  const int ny = det ? g.n1y : g.n0y;
  const int nz = det ? g.n1z : g.n0z;
  const float p0 = det ? g.pos1x : g.pos0x;
  const float p1 = ((float)idy - 0.5f * (float)(ny - 1)) * e1;
  const float p2 = ((float)idz - 0.5f * (float)(nz - 1)) * e2;
  const float sn = g.sin_[ida], cs = g.cos_[ida];
  const float t0 = cs * e0, t1 = 0.0f, t2 = sn * e2;
  const float t3 = cs * (p0 - 0.5f * e0) + sn * (p2 - 0.5f * e2);
  const float t4 = 0.0f, t5 = e1, t6 = 0.0f, t7 = p1 - 0.5f * e1;
  const float t8 = (-sn) * e0, t9 = 0.0f, t10 = cs * e2;
  const float t11 = (-sn) * (p0 - 0.5f * e0) + cs * (p2 - 0.5f * e2);
  out[0] = ((t0 * b0 + t1 * b1) + t2 * b2) + t3 * 1.0f;
  out[1] = ((t4 * b0 + t5 * b1) + t6 * b2) + t7 * 1.0f;
  out[2] = ((t8 * b0 + t9 * b1) + t10 * b2) + t11 * 1.0f;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; round++) {
    const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0;
    const uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Uniform in (0,1): odd multiples of 2^-24 (exact in float).
__device__ __forceinline__ float u01(uint32_t x) {
  return (float)((x >> 8) | 1u) * 5.9604644775390625e-08f;
}

// Ray r of channel cnl: start in the detector-0 segment, end in detector-1.
__device__ __forceinline__ void make_ray(const Geo& g, const ChannelIds& ci, int cnl, int r,
                                         float ray[6]) {
  uint32_t a[4] = {(uint32_t)r, (uint32_t)cnl, 0u, 0u};
  uint32_t b[4] = {(uint32_t)r, (uint32_t)cnl, 1u, 0u};
  philox4x32_10(a, g.key0, g.key1);
  philox4x32_10(b, g.key0, g.key1);
  trafo(g, 0, u01(a[0]), u01(a[1]), u01(a[2]), ci.id0z, ci.id0y, ci.ida, &ray[0]);
  trafo(g, 1, u01(a[3]), u01(b[0]), u01(b[1]), ci.id1z, ci.id1y, ci.ida, &ray[3]);
}

// |a| (distancePointLine.h:19-24).
__device__ __forceinline__ float absolute3(float a0, float a1, float a2) {
  return sqrtf((a0 * a0 + a1 * a1) + a2 * a2);
}

// distance of p from the line through a and b (distancePointLine.h:41-55).
__device__ __forceinline__ float distance_pl(const float a[3], const float b[3],
                                             const float p[3]) {
  float ap[3], ab[3], c[3];
#pragma unroll
  for (int d = 0; d < 3; d++) { ap[d] = p[d] - a[d]; ab[d] = b[d] - a[d]; }
#pragma unroll
  for (int d = 0; d < 3; d++)
    c[d] = ab[(d + 1) % 3] * ap[(d + 2) % 3] - ab[(d + 2) % 3] * ap[(d + 1) % 3];
  return absolute3(c[0], c[1], c[2]) / absolute3(ab[0], ab[1], ab[2]);
}

// Pixel centres of a channel (getWorkqueue.hpp:64-69, test() of
// getSystemMatrixDeviceOnly.cu:60-65).
__device__ __forceinline__ void pixel_centres(const Geo& g, const ChannelIds& ci,
                                              float p0[3], float p1[3]) {
  trafo(g, 0, 0.5f, 0.5f, 0.5f, ci.id0z, ci.id0y, ci.ida, p0);
  trafo(g, 1, 0.5f, 0.5f, 0.5f, ci.id1z, ci.id1y, ci.ida, p1);
}

// Voxel ids (VoxelGridLinIndex.hpp:220-282).
__device__ __forceinline__ void voxel_ids(const Geo& g, int v, int& ix, int& iy, int& iz) {
  const int nyz = g.gn[2] * g.gn[1];
  ix = v / nyz;
  const int t = v % nyz;
  iy = t / g.gn[2];
  iz = t % g.gn[2];
}

// Candidate predicate (getWorkqueue.hpp:72-78).
__device__ __forceinline__ bool is_candidate(const Geo& g, const float p0[3],
                                             const float p1[3], int v) {
  int ix, iy, iz;
  voxel_ids(g, v, ix, iy, iz);
  float c[3];
  c[0] = g.go[0] + ((float)ix + 0.5f) * g.gd[0];
  c[1] = g.go[1] + ((float)iy + 0.5f) * g.gd[1];
  c[2] = g.go[2] + ((float)iz + 0.5f) * g.gd[2];
  return distance_pl(p0, p1, c) < g.sumRadii;
}

// Voxel box corners (getSystemMatrixDeviceOnly.cu:166-175).
__device__ __forceinline__ void voxel_box(const Geo& g, int v, float box[6]) {
  int id[3];
  voxel_ids(g, v, id[0], id[1], id[2]);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    box[d] = g.go[d] + (float)id[d] * g.gd[d];
    box[3 + d] = g.go[d] + (float)(id[d] + 1) * g.gd[d];
  }
}

// intersectionLength for a ray with start p, direction den = end - start and
// length len = |den| (getSystemMatrixDeviceOnly.cu:95-132), general masked
// form: axes with den == 0 are ignored, as in the reference.
__device__ __forceinline__ float slab_length_masked(const float box[6], const float p[3],
                                                    const float den[3], float len) {
  bool good = false;
  float lo = 0.0f, hi = 0.0f;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (den[d] != 0.0f) {
      float a = (box[d] - p[d]) / den[d];
      float b = (box[3 + d] - p[d]) / den[d];
      if (b < a) { float t = a; a = b; b = t; }
      if (!good) { lo = a; hi = b; good = true; }
      else { lo = fmaxf(lo, a); hi = fminf(hi, b); }
    }
  }
  if (!good) return 0.0f;
  if (!(lo < hi)) return 0.0f;
  return (hi - lo) * len;
}

// Same for a ray with den nonzero on all three axes.
__device__ __forceinline__ float slab_length(const float box[6], float p0, float p1, float p2,
                                             float d0, float d1, float d2, float len) {
  float a0 = (box[0] - p0) / d0, b0 = (box[3] - p0) / d0;
  float a1 = (box[1] - p1) / d1, b1 = (box[4] - p1) / d1;
  float a2 = (box[2] - p2) / d2, b2 = (box[5] - p2) / d2;
  if (b0 < a0) { float t = a0; a0 = b0; b0 = t; }
  if (b1 < a1) { float t = a1; a1 = b1; b1 = t; }
  if (b2 < a2) { float t = a2; a2 = b2; b2 = t; }
  const float lo = fmaxf(fmaxf(a0, a1), a2);
  const float hi = fminf(fminf(b0, b1), b2);
  if (!(lo < hi)) return 0.0f;
  return (hi - lo) * len;
}

}  // namespace epet
