/*
 * epet_oracle.c -- CPU restatement of the epetreco MLEM hot path.
 * TEST INFRASTRUCTURE ONLY (see epet_oracle.h).
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math -fopenmp (oracle/Makefile).
 * The no-contraction flag keeps every float product rounded on its own, as the
 * reference's host code is on x86-64 without FMA.
 */
#include "epet_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* index maps: MeasurementSetupLinIndex.hpp:32-165                          */
/* ------------------------------------------------------------------------ */
int eo_ms_linid(const eo_setup* s, int id0z, int id0y, int id1z, int id1y, int ida) {
  /* MeasurementSetupLinIndex.hpp:40-44 */
  return ida  * (s->n1y * s->n1z * s->n0y * s->n0z)
       + id0z * (s->n1y * s->n1z * s->n0y)
       + id0y * (s->n1y * s->n1z)
       + id1z * (s->n1y)
       + id1y;
}
int eo_ms_ida(const eo_setup* s, int lin) {          /* :65 */
  return lin / (s->n1y * s->n1z * s->n0y * s->n0z);
}
int eo_ms_id0z(const eo_setup* s, int lin) {         /* :86-88 */
  int t = lin % (s->n1y * s->n1z * s->n0y * s->n0z);
  return t / (s->n1y * s->n1z * s->n0y);
}
int eo_ms_id0y(const eo_setup* s, int lin) {         /* :109-112 */
  int t = lin % (s->n1y * s->n1z * s->n0y * s->n0z);
  t %= (s->n1y * s->n1z * s->n0y);
  return t / (s->n1y * s->n1z);
}
int eo_ms_id1z(const eo_setup* s, int lin) {         /* :133-137 */
  int t = lin % (s->n1y * s->n1z * s->n0y * s->n0z);
  t %= (s->n1y * s->n1z * s->n0y);
  t %= (s->n1y * s->n1z);
  return t / s->n1y;
}
int eo_ms_id1y(const eo_setup* s, int lin) {         /* :158-163 */
  int t = lin % (s->n1y * s->n1z * s->n0y * s->n0z);
  t %= (s->n1y * s->n1z * s->n0y);
  t %= (s->n1y * s->n1z);
  return t % s->n1y;
}

/* VoxelGridLinIndex.hpp:193-282 -- z fastest */
int eo_vg_linid(const eo_grid* g, int idx, int idy, int idz) {   /* :201-203 */
  return idx * (g->n[2] * g->n[1]) + idy * g->n[2] + idz;
}
int eo_vg_idx(const eo_grid* g, int lin) { return lin / (g->n[2] * g->n[1]); }  /* :227 */
int eo_vg_idy(const eo_grid* g, int lin) {                                       /* :251-253 */
  int t = lin % (g->n[2] * g->n[1]);
  return t / g->n[2];
}
int eo_vg_idz(const eo_grid* g, int lin) {                                       /* :277-280 */
  int t = lin % (g->n[2] * g->n[1]);
  return t % g->n[2];
}

/* ------------------------------------------------------------------------ */
/* geometry                                                                 */
/* ------------------------------------------------------------------------ */
void eo_angle_sincos(const eo_setup* s, int ida, float* sin_, float* cos_) {
  /* MeasurementSetupTrafo2CartCoord.hpp:67-68: T(M_PI)*ida*(da)/T(180.0) */
  float arg = (((float)M_PI * (float)ida) * s->da) / 180.0f;
  *sin_ = sinf(arg);
  *cos_ = cosf(arg);
}

void eo_trafo(const eo_setup* s, int det, const float box[3],
              int idz, int idy, int ida, float out[3]) {
  /* MeasurementSetupTrafo2CartCoord.hpp:56-104 (det 0) / :225-273 (det 1) */
  const float e0 = s->segx, e1 = s->segy, e2 = s->segz;
  const int ny = det ? s->n1y : s->n0y;
  const int nz = det ? s->n1z : s->n0z;
  float p0 = det ? s->pos1x : s->pos0x;
  float p1 = ((float)idy - 0.5f * (float)(ny - 1)) * e1;
  float p2 = ((float)idz - 0.5f * (float)(nz - 1)) * e2;
  float sn, cs;
  eo_angle_sincos(s, ida, &sn, &cs);
  float t[12];
  t[0]  = cs * e0;
  t[1]  = 0.0f;
  t[2]  = sn * e2;
  t[3]  = cs * (p0 - 0.5f * e0) + sn * (p2 - 0.5f * e2);
  t[4]  = 0.0f;
  t[5]  = e1;
  t[6]  = 0.0f;
  t[7]  = p1 - 0.5f * e1;
  t[8]  = (-sn) * e0;
  t[9]  = 0.0f;
  t[10] = cs * e2;
  t[11] = (-sn) * (p0 - 0.5f * e0) + cs * (p2 - 0.5f * e2);
  const float b0 = box[0], b1 = box[1], b2 = box[2];
  out[0] = ((t[0] * b0 + t[1] * b1) + t[2] * b2) + t[3] * 1.0f;
  out[1] = ((t[4] * b0 + t[5] * b1) + t[6] * b2) + t[7] * 1.0f;
  out[2] = ((t[8] * b0 + t[9] * b1) + t[10] * b2) + t[11] * 1.0f;
}

/* distancePointLine.h:19-55 */
float eo_absolute(const float a[3]) {
  return sqrtf((a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]);
}
void eo_cross(float r[3], const float a[3], const float b[3]) {
  for (int d = 0; d < 3; d++)
    r[d] = a[(d + 1) % 3] * b[(d + 2) % 3] - a[(d + 2) % 3] * b[(d + 1) % 3];
}
float eo_distance(const float a[3], const float b[3], const float p[3]) {
  float ap[3], ab[3], c[3];
  for (int d = 0; d < 3; d++) { ap[d] = p[d] - a[d]; ab[d] = b[d] - a[d]; }
  eo_cross(c, ab, ap);
  return eo_absolute(c) / eo_absolute(ab);
}

/* ------------------------------------------------------------------------ */
/* workqueue: getWorkqueue.hpp:33-223                                       */
/* ------------------------------------------------------------------------ */
static float sum_radii(const eo_setup* s, const eo_grid* g) {
  /* getWorkqueue.hpp:54-56: 0.5 (double) * (float + float), narrowed */
  float ve[3] = {g->d[0], g->d[1], g->d[2]};
  float pe[3] = {s->segx, s->segy, s->segz};
  return (float)(0.5 * (double)(eo_absolute(ve) + eo_absolute(pe)));
}

static void pixel_centres(const eo_setup* s, int cnl, float p0[3], float p1[3]) {
  static const float centre[3] = {0.5f, 0.5f, 0.5f};
  int ida = eo_ms_ida(s, cnl);
  eo_trafo(s, 0, centre, eo_ms_id0z(s, cnl), eo_ms_id0y(s, cnl), ida, p0);
  eo_trafo(s, 1, centre, eo_ms_id1z(s, cnl), eo_ms_id1y(s, cnl), ida, p1);
}

static void voxel_centre(const eo_grid* g, int vxl, float c[3]) {
  /* getWorkqueue.hpp:72-74: O + (id + 0.5) * D */
  c[0] = g->o[0] + ((float)eo_vg_idx(g, vxl) + 0.5f) * g->d[0];
  c[1] = g->o[1] + ((float)eo_vg_idy(g, vxl) + 0.5f) * g->d[1];
  c[2] = g->o[2] + ((float)eo_vg_idz(g, vxl) + 0.5f) * g->d[2];
}

int eo_is_candidate(const eo_setup* s, const eo_grid* g, int cnl, int vxl) {
  float p0[3], p1[3], c[3];
  pixel_centres(s, cnl, p0, p1);
  voxel_centre(g, vxl, c);
  return eo_distance(p0, p1, c) < sum_radii(s, g);   /* :78 */
}

long long eo_workqueue(const eo_setup* s, const eo_grid* g,
                       const int* list, int list_size,
                       int* listId, int* vxlId, long long cap,
                       int* out_cnl, int* out_vxl) {
  const int V = g->n[0] * g->n[1] * g->n[2];
  const float sr = sum_radii(s, g);
  long long found = 0;
  int li = *listId, vi = *vxlId;
  int cached = -1;
  float p0[3], p1[3];
  while (li < list_size && (cap < 0 || found < cap)) {
    int cnl = list[li];
    if (li != cached) { pixel_centres(s, cnl, p0, p1); cached = li; }
    float c[3];
    voxel_centre(g, vi, c);
    int hit = eo_distance(p0, p1, c) < sr;
    if (hit) {
      if (out_cnl) out_cnl[found] = cnl;
      if (out_vxl) out_vxl[found] = vi;
      found++;
    }
    /* cursor increment (getWorkqueue.hpp:83-92, 99-108) */
    if (vi < V - 1) vi++;
    else { vi = 0; li++; }
  }
  *listId = li;
  *vxlId = vi;
  return found;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 and rays                                                   */
/* ------------------------------------------------------------------------ */
void eo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; round++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

float eo_u01(uint32_t x) {
  return (float)((x >> 8) | 1u) * 5.9604644775390625e-08f;  /* 2^-24 */
}

void eo_ray(const eo_setup* s, int cnl, int r, uint64_t seed, float ray[6]) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t c0[4] = {(uint32_t)r, (uint32_t)cnl, 0u, 0u};
  uint32_t c1[4] = {(uint32_t)r, (uint32_t)cnl, 1u, 0u};
  uint32_t a[4], b[4];
  eo_philox4x32_10(c0, key, a);
  eo_philox4x32_10(c1, key, b);
  float u0[3] = {eo_u01(a[0]), eo_u01(a[1]), eo_u01(a[2])};
  float u1[3] = {eo_u01(a[3]), eo_u01(b[0]), eo_u01(b[1])};
  int ida = eo_ms_ida(s, cnl);
  eo_trafo(s, 0, u0, eo_ms_id0z(s, cnl), eo_ms_id0y(s, cnl), ida, &ray[0]);
  eo_trafo(s, 1, u1, eo_ms_id1z(s, cnl), eo_ms_id1y(s, cnl), ida, &ray[3]);
}

void eo_ray_center(const eo_setup* s, int cnl, float ray[6]) {
  pixel_centres(s, cnl, &ray[0], &ray[3]);
}

/* ------------------------------------------------------------------------ */
/* intersection length and system-matrix values                             */
/* ------------------------------------------------------------------------ */
float eo_intersection_length(const float vxl[6], const float ray[6]) {
  /* getSystemMatrixDeviceOnly.cu:95-132 */
  int sect[3];
  for (int d = 0; d < 3; d++) sect[d] = ray[d] != ray[3 + d];
  if (!(sect[0] || sect[1] || sect[2])) return 0.0f;
  float amin[3] = {0, 0, 0}, amax[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) {
    if (sect[d]) {
      float den = ray[3 + d] - ray[d];
      amin[d] = (vxl[d] - ray[d]) / den;
      amax[d] = (vxl[3 + d] - ray[d]) / den;
      if (amax[d] < amin[d]) { float t = amin[d]; amin[d] = amax[d]; amax[d] = t; }
    }
  }
  /* MaxFunctor<3> / MinFunctor<3> over the masked entries
   * (ChordsCalc_lowlevel.hpp:30-156): value of the first true entry, then
   * max/min with every further true entry. */
  int good = 0;
  float lo = 0.0f, hi = 0.0f;
  for (int d = 0; d < 3; d++) {
    if (!sect[d]) continue;
    if (!good) { lo = amin[d]; hi = amax[d]; good = 1; }
    else { lo = fmaxf(lo, amin[d]); hi = fminf(hi, amax[d]); }
  }
  if (!(lo < hi)) return 0.0f;
  float dx = ray[3] - ray[0], dy = ray[4] - ray[1], dz = ray[5] - ray[2];
  return (hi - lo) * sqrtf((dx * dx + dy * dy) + dz * dz);
}

void eo_voxel_box(const eo_grid* g, int vxl, float box[6]) {
  /* getSystemMatrixDeviceOnly.cu:166-175 */
  int id[3] = {eo_vg_idx(g, vxl), eo_vg_idy(g, vxl), eo_vg_idz(g, vxl)};
  for (int d = 0; d < 3; d++) {
    box[d]     = g->o[d] + (float)id[d] * g->d[d];
    box[3 + d] = g->o[d] + (float)(id[d] + 1) * g->d[d];
  }
}

float eo_calc_sme_rays(const eo_grid* g, int vxl, const float* rays, int nrays) {
  float box[6];
  eo_voxel_box(g, vxl, box);
  float a = 0.0f;
  for (int r = 0; r < nrays; r++) a += eo_intersection_length(box, rays + 6 * r);
  return a / (float)nrays;   /* getSystemMatrixDeviceOnly.cu:199 */
}

float eo_calc_sme(const eo_setup* s, const eo_grid* g, int cnl, int vxl,
                  int nrays, uint64_t seed) {
  float* rays = (float*)malloc(sizeof(float) * 6 * (size_t)nrays);
  for (int r = 0; r < nrays; r++) eo_ray(s, cnl, r, seed, rays + 6 * r);
  float a = eo_calc_sme_rays(g, vxl, rays, nrays);
  free(rays);
  return a;
}

void eo_sm_from_workqueue(const eo_setup* s, const eo_grid* g,
                          const int* cnl, const int* vxl, float* val,
                          long long n, int nrays, uint64_t seed) {
#pragma omp parallel
  {
    float* rays = (float*)malloc(sizeof(float) * 6 * (size_t)nrays);
    int have = -1;
#pragma omp for schedule(static)
    for (long long i = 0; i < n; i++) {
      if (cnl[i] != have) {
        for (int r = 0; r < nrays; r++) eo_ray(s, cnl[i], r, seed, rays + 6 * r);
        have = cnl[i];
      }
      val[i] = eo_calc_sme_rays(g, vxl[i], rays, nrays);
    }
    free(rays);
  }
}

long long eo_sm_rows(const eo_setup* s, const eo_grid* g, const int* list, int m,
                     int nrays, uint64_t seed, int keep_zeros,
                     int64_t* rowptr, int* col, float* val, long long cap) {
  const int V = g->n[0] * g->n[1] * g->n[2];
  const float sr = sum_radii(s, g);
  int64_t* cnt = (int64_t*)calloc((size_t)m + 1, sizeof(int64_t));
  /* pass 1: per-row counts */
#pragma omp parallel
  {
    float* rays = (float*)malloc(sizeof(float) * 6 * (size_t)(nrays > 0 ? nrays : 1));
#pragma omp for schedule(dynamic, 4)
    for (int i = 0; i < m; i++) {
      int cnl = list[i];
      float p0[3], p1[3];
      pixel_centres(s, cnl, p0, p1);
      for (int r = 0; r < nrays; r++) eo_ray(s, cnl, r, seed, rays + 6 * r);
      int64_t k = 0;
      for (int v = 0; v < V; v++) {
        float c[3];
        voxel_centre(g, v, c);
        if (!(eo_distance(p0, p1, c) < sr)) continue;
        if (keep_zeros || eo_calc_sme_rays(g, v, rays, nrays) != 0.0f) k++;
      }
      cnt[i + 1] = k;
    }
    free(rays);
  }
  for (int i = 0; i < m; i++) cnt[i + 1] += cnt[i];
  long long nnz = cnt[m];
  if (rowptr) memcpy(rowptr, cnt, sizeof(int64_t) * ((size_t)m + 1));
  if (col == NULL || val == NULL) { free(cnt); return nnz; }
  if (cap >= 0 && nnz > cap) { free(cnt); return -1; }
  /* pass 2: fill */
#pragma omp parallel
  {
    float* rays = (float*)malloc(sizeof(float) * 6 * (size_t)(nrays > 0 ? nrays : 1));
#pragma omp for schedule(dynamic, 4)
    for (int i = 0; i < m; i++) {
      int cnl = list[i];
      float p0[3], p1[3];
      pixel_centres(s, cnl, p0, p1);
      for (int r = 0; r < nrays; r++) eo_ray(s, cnl, r, seed, rays + 6 * r);
      int64_t k = cnt[i];
      for (int v = 0; v < V; v++) {
        float c[3];
        voxel_centre(g, v, c);
        if (!(eo_distance(p0, p1, c) < sr)) continue;
        float a = eo_calc_sme_rays(g, v, rays, nrays);
        if (keep_zeros || a != 0.0f) { col[k] = v; val[k] = a; k++; }
      }
    }
    free(rays);
  }
  free(cnt);
  return nnz;
}

/* ------------------------------------------------------------------------ */
/* Siddon chords: ChordsCalc_lowlevel.hpp:173-654                           */
/* ------------------------------------------------------------------------ */
/* masked max / min of 3 (MaxFunctor<3>, MinFunctor<3>, :30-156) */
static void masked_max3(float* out, int* good, const float a[3], const int b[3]) {
  int g = 0; float v = 0.0f;
  for (int d = 0; d < 3; d++) {
    if (!b[d]) continue;
    if (!g) { v = a[d]; g = 1; } else v = fmaxf(v, a[d]);
  }
  *out = v; *good = g;
}
static void masked_min3(float* out, int* good, const float a[3], const int b[3]) {
  int g = 0; float v = 0.0f;
  for (int d = 0; d < 3; d++) {
    if (!b[d]) continue;
    if (!g) { v = a[d]; g = 1; } else v = fminf(v, a[d]);
  }
  *out = v; *good = g;
}
static float alpha_from_id(int i, int d, const float ray[6], const float O[3],
                           const float D[3]) {
  /* :421-434 */
  float divisor = ray[d + 3] - ray[d];
  divisor += (float)(divisor == 0.0f);
  return (O[d] + (float)i * D[d] - ray[d]) / divisor;
}
static float phi_from_alpha(float alpha, int d, const float ray[6], const float O[3],
                            const float D[3]) {
  /* :453-465 */
  return (ray[d] + alpha * (ray[d + 3] - ray[d]) - O[d]) / D[d];
}

int eo_get_chords(float* chords, int* voxelIds, int nChords, const float ray[6],
                  const float O[3], const float D[3], const int N[3]) {
  float admin[3], admax[3];
  int crosses[3];
  for (int d = 0; d < 3; d++) {        /* getAlphaDimmin/max :484-533 */
    float a0 = alpha_from_id(0, d, ray, O, D), a1 = alpha_from_id(N[d], d, ray, O, D);
    admin[d] = fminf(a0, a1);
    admax[d] = fmaxf(a0, a1);
  }
  /* getCrossesPlanes :373-399 */
  int goodPar = 1;
  for (int d = 0; d < 3; d++) {
    crosses[d] = (ray[d + 3] - ray[d] != 0.0f);
    goodPar &= (crosses[d] ||
                (!crosses[d] && ray[d] >= O[d] && ray[d] < O[d] + D[d] * (float)N[d]));
  }
  for (int d = 0; d < 3; d++) crosses[d] &= goodPar;
  float aMin, aMax; int aMinGood, aMaxGood;
  masked_max3(&aMin, &aMinGood, admin, crosses);     /* getAlphaMin */
  masked_min3(&aMax, &aMaxGood, admax, crosses);     /* getAlphaMax */
  aMinGood &= ((double)aMin >= 0. && (double)aMin <= 1.);   /* :208-209 */
  aMaxGood &= ((double)aMax >= 0. && (double)aMax <= 1.);
  if (aMin > aMax || !aMinGood || !aMaxGood) return 0;      /* :218 */
  float aup[3]; int idup[3];
  for (int d = 0; d < 3; d++) {        /* getAlphaDimup :612-629 */
    float dv = ray[d + 3] - ray[d];
    float divisor = sqrtf(dv * dv);
    divisor += (float)(int)(divisor == 0.0f);
    aup[d] = D[d] / divisor;
  }
  for (int d = 0; d < 3; d++)          /* getIdDimup :642-654 */
    idup[d] = (int)(ray[d] < ray[d + 3]) - (int)(ray[d] >= ray[d + 3]);
  float anext[3];
  for (int d = 0; d < 3; d++) anext[d] = admin[d] + aup[d];
  int id[3];
  float aNext; int aNextExists;
  masked_min3(&aNext, &aNextExists, anext, crosses);
  for (int d = 0; d < 3; d++)
    id[d] = (int)floorf(phi_from_alpha(0.5f * (aMin + aNext), d, ray, O, D));
  float aCurr = aMin;
  float dx = ray[3] - ray[0], dy = ray[4] - ray[1], dz = ray[5] - ray[2];
  const float length = sqrtf((dx * dx + dy * dy) + dz * dz);   /* getLength :588-599 */
  int k = 0;
  while (aCurr < aMax) {               /* :292-356 */
    masked_min3(&aNext, &aNextExists, anext, crosses);
    for (int d = 0; d < 3; d++) {
      int crossed = (anext[d] == aNext);
      if (crossed) {
        if (k >= nChords) return -1;
        chords[k] = 1.0f * (anext[d] - aCurr) * length;
        for (int w = 0; w < 3; w++) voxelIds[3 * k + w] = id[w];
        k++;
        aCurr = anext[d];
        anext[d] += aup[d];
        id[d] += idup[d];
      }
    }
  }
  return k;
}

/* ------------------------------------------------------------------------ */
/* sparse formats                                                           */
/* ------------------------------------------------------------------------ */
typedef struct { int row, col; float val; long long pos; } coo_entry;
static int coo_less(const void* pa, const void* pb) {
  const coo_entry* a = (const coo_entry*)pa;
  const coo_entry* b = (const coo_entry*)pb;
  if (a->row != b->row) return a->row < b->row ? -1 : 1;   /* cooSort.hpp:34-43 */
  if (a->col != b->col) return a->col < b->col ? -1 : 1;
  return a->pos < b->pos ? -1 : (a->pos > b->pos);          /* stable */
}
void eo_coo_sort(float* val, int* row, int* col, long long n) {
  coo_entry* e = (coo_entry*)malloc(sizeof(coo_entry) * (size_t)(n > 0 ? n : 1));
  for (long long i = 0; i < n; i++) { e[i].row = row[i]; e[i].col = col[i]; e[i].val = val[i]; e[i].pos = i; }
  qsort(e, (size_t)n, sizeof(coo_entry), coo_less);
  for (long long i = 0; i < n; i++) { row[i] = e[i].row; col[i] = e[i].col; val[i] = e[i].val; }
  free(e);
}

void eo_coo2csr(int* csrRowPtr, const int* cooRowId, long long nnz, int m) {
  /* row pointer = number of entries with row < r (zero-based) */
  for (int r = 0; r <= m; r++) csrRowPtr[r] = 0;
  for (long long i = 0; i < nnz; i++) csrRowPtr[cooRowId[i] + 1]++;
  for (int r = 0; r < m; r++) csrRowPtr[r + 1] += csrRowPtr[r];
}

void eo_csr2ecsr(int* ecsr, const int* yRowId, int nnzr, const int* csr, int m) {
  /* convertCsr2Ecsr.hpp:43-46 */
  for (int i = 0; i < nnzr; i++) ecsr[i] = csr[yRowId[i]];
  ecsr[nnzr] = csr[m];
}

/* ------------------------------------------------------------------------ */
/* algebra                                                                  */
/* ------------------------------------------------------------------------ */
void eo_csrmv(int trans, int m, int n, float alpha, const float* val,
              const int64_t* rowptr, const int* col, const float* x,
              float beta, float* y) {
  if (!trans) {
    for (int i = 0; i < m; i++) {
      double acc = 0.0;
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) acc += (double)val[k] * x[col[k]];
      y[i] = (float)((double)alpha * acc + (beta == 0.0f ? 0.0 : (double)beta * y[i]));
    }
  } else {
    double* acc = (double*)calloc((size_t)n, sizeof(double));
    for (int i = 0; i < m; i++)
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) acc[col[k]] += (double)val[k] * x[i];
    for (int j = 0; j < n; j++)
      y[j] = (float)((double)alpha * acc[j] + (beta == 0.0f ? 0.0 : (double)beta * y[j]));
    free(acc);
  }
}

/* mlemOperations.hpp:22-139 */
void eo_divides(float* res, const float* a, const float* b, long long n) {
  for (long long i = 0; i < n; i++) res[i] = a[i] / b[i];
}
void eo_divides_multiplies(float* res, const float* a, const float* b,
                           const float* c, long long n) {
  for (long long i = 0; i < n; i++) {
    float dvs = c[i];
    dvs += (float)(dvs == 0.0f);   /* :48-50 */
    res[i] = a[i] * b[i] / dvs;
  }
}
void eo_scales(float* x, float alpha, long long n) {
  for (long long i = 0; i < n; i++) x[i] = alpha * x[i];
}
float eo_sum(const float* x, long long n) {
  float s = 0.0f;
  for (long long i = 0; i < n; i++) s += x[i];
  return s;
}
double eo_sum_f64(const float* x, long long n) {
  double s = 0.0;
  for (long long i = 0; i < n; i++) s += x[i];
  return s;
}

/* ------------------------------------------------------------------------ */
/* MLEM: tests/reco.cu:209-300                                              */
/* ------------------------------------------------------------------------ */
void eo_mlem_f32(int m, int n, const float* val, const int64_t* rowptr,
                 const int* col, const float* y, const float* s_in, float* x,
                 int niter) {
  float* s = (float*)malloc(sizeof(float) * (size_t)n);
  float* c = (float*)malloc(sizeof(float) * (size_t)n);
  float* yt = (float*)malloc(sizeof(float) * (size_t)(m > 0 ? m : 1));
  float* e = (float*)malloc(sizeof(float) * (size_t)(m > 0 ? m : 1));
  memcpy(s, s_in, sizeof(float) * (size_t)n);
  {
    float norm = eo_sum(s, n);                      /* reco.cu:182-183 */
    eo_scales(s, (float)(1. / norm), n);
  }
  for (int it = 0; it < niter; it++) {
    for (int j = 0; j < n; j++) c[j] = 0.0f;        /* :212-213 */
    for (int i = 0; i < m; i++) {                   /* :236-238, beta = 0 */
      float acc = 0.0f;
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) acc += val[k] * x[col[k]];
      yt[i] = acc;
    }
    eo_divides(e, y, yt, m);                        /* :242-243 */
    for (int i = 0; i < m; i++)                     /* :247-249, beta = 1 */
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) c[col[k]] += val[k] * e[i];
    eo_divides_multiplies(x, x, c, s, n);           /* :273, :277 */
    float norm = eo_sum(x, n);                      /* :280-281 */
    eo_scales(x, (float)(1. / norm), n);
  }
  free(s); free(c); free(yt); free(e);
}

void eo_mlem_f64(int m, int n, const float* val, const int64_t* rowptr,
                 const int* col, const float* y, const float* s_in, double* x,
                 int niter) {
  double* s = (double*)malloc(sizeof(double) * (size_t)n);
  double* c = (double*)malloc(sizeof(double) * (size_t)n);
  double* e = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double ss = 0.0;
  for (int j = 0; j < n; j++) ss += s_in[j];
  for (int j = 0; j < n; j++) s[j] = s_in[j] / ss;
  for (int it = 0; it < niter; it++) {
    for (int j = 0; j < n; j++) c[j] = 0.0;
    for (int i = 0; i < m; i++) {
      double acc = 0.0;
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) acc += (double)val[k] * x[col[k]];
      e[i] = acc > 0.0 ? (double)y[i] / acc : 0.0;
    }
    for (int i = 0; i < m; i++)
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) c[col[k]] += (double)val[k] * e[i];
    double sum = 0.0;
    for (int j = 0; j < n; j++) {
      double dvs = s[j] + (s[j] == 0.0 ? 1.0 : 0.0);
      x[j] = x[j] * c[j] / dvs;
      sum += x[j];
    }
    for (int j = 0; j < n; j++) x[j] /= sum;
  }
  free(s); free(c); free(e);
}

void eo_mlem_f64_frames(int m, int n, int F, const float* val, const int64_t* rowptr,
                        const int* col, const float* Y, const float* s_in, double* X,
                        const int* niter_f, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  const int64_t nnz = rowptr[m];
  /* transpose by counting sort; rows of the transpose list channels ascending */
  int64_t* tptr = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  int* trow = (int*)malloc(sizeof(int) * (size_t)(nnz > 0 ? nnz : 1));
  float* tval = (float*)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
  for (int64_t k = 0; k < nnz; k++) tptr[col[k] + 1]++;
  for (int j = 0; j < n; j++) tptr[j + 1] += tptr[j];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < n; j++) fill[j] = tptr[j];
  for (int i = 0; i < m; i++)
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
      const int64_t d = fill[col[k]]++;
      trow[d] = i;
      tval[d] = val[k];
    }
  free(fill);
  double* s = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double ss = 0.0;
  for (int j = 0; j < n; j++) ss += s_in[j];
  for (int j = 0; j < n; j++) s[j] = s_in[j] / ss;
  double* E = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1) * (size_t)F);
  double* C = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1) * (size_t)F);
  int maxit = 0;
  for (int f = 0; f < F; f++) maxit = niter_f[f] > maxit ? niter_f[f] : maxit;
  for (int it = 0; it < maxit; it++) {
    /* frames in blocks of 64, entries ascending within each frame's sum */
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int i = 0; i < m; i++)
      for (int f0 = 0; f0 < F; f0 += 64) {
        const int nf = F - f0 < 64 ? F - f0 : 64;
        double acc[64];
        for (int f = 0; f < nf; f++) acc[f] = 0.0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
          const double v = (double)val[k];
          const double* xr = X + (size_t)col[k] * F + f0;
          for (int f = 0; f < nf; f++) acc[f] += v * xr[f];
        }
        for (int f = 0; f < nf; f++) {
          const size_t q = (size_t)i * F + f0 + f;
          E[q] = acc[f] > 0.0 ? (double)Y[q] / acc[f] : 0.0;
        }
      }
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int j = 0; j < n; j++)
      for (int f0 = 0; f0 < F; f0 += 64) {
        const int nf = F - f0 < 64 ? F - f0 : 64;
        double acc[64];
        for (int f = 0; f < nf; f++) acc[f] = 0.0;
        for (int64_t k = tptr[j]; k < tptr[j + 1]; k++) {
          const double v = (double)tval[k];
          const double* er = E + (size_t)trow[k] * F + f0;
          for (int f = 0; f < nf; f++) acc[f] += v * er[f];
        }
        for (int f = 0; f < nf; f++) C[(size_t)j * F + f0 + f] = acc[f];
      }
    for (int f = 0; f < F; f++) {
      if (it >= niter_f[f]) continue;
      double sum = 0.0;
      for (int j = 0; j < n; j++) {
        const double dvs = s[j] + (s[j] == 0.0 ? 1.0 : 0.0);
        double* xj = &X[(size_t)j * F + f];
        *xj = *xj * C[(size_t)j * F + f] / dvs;
        sum += *xj;
      }
      for (int j = 0; j < n; j++) X[(size_t)j * F + f] /= sum;
    }
  }
  free(tptr); free(trow); free(tval); free(s); free(E); free(C);
}

/* Static row range [*r0, *r1) of thread t of nt (contiguous, in row order). */
static void row_range(int m, int t, int nt, int* r0, int* r1) {
  *r0 = (int)(((int64_t)m * t) / nt);
  *r1 = (int)(((int64_t)m * (t + 1)) / nt);
}

/* tests/reco.cu:209-300 loop for F frames (as eo_mlem_f64_frames), with the
 * back projection c = A^T e scattered into per-thread FP64 images. */
void eo_mlem_f64_frames_par(int m, int n, int F, const float* val, const int64_t* rowptr,
                            const int* col, const float* Y, const float* s_in, double* X,
                            const int* niter_f, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  const size_t nF = (size_t)(n > 0 ? n : 1) * (size_t)F;
  double* s = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double ss = 0.0;
  for (int j = 0; j < n; j++) ss += s_in[j];
  for (int j = 0; j < n; j++) s[j] = s_in[j] / ss;
  double* E = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1) * (size_t)F);
  double* C = (double*)malloc(sizeof(double) * nF);
  double* priv = (double*)malloc(sizeof(double) * nF * (size_t)nthreads);
  int maxit = 0;
  for (int f = 0; f < F; f++) maxit = niter_f[f] > maxit ? niter_f[f] : maxit;
  for (int it = 0; it < maxit; it++) {
    /* forward projection and ratio e = y / (A x), 0 where A x = 0 */
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
    for (int i = 0; i < m; i++)
      for (int f0 = 0; f0 < F; f0 += 64) {
        const int nf = F - f0 < 64 ? F - f0 : 64;
        double acc[64];
        for (int f = 0; f < nf; f++) acc[f] = 0.0;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
          const double v = (double)val[k];
          const double* xr = X + (size_t)col[k] * F + f0;
          for (int f = 0; f < nf; f++) acc[f] += v * xr[f];
        }
        for (int f = 0; f < nf; f++) {
          const size_t q = (size_t)i * F + f0 + f;
          E[q] = acc[f] > 0.0 ? (double)Y[q] / acc[f] : 0.0;
        }
      }
    /* back projection: thread t scatters its rows into priv[t], then the
     * images are summed in thread order */
#pragma omp parallel num_threads(nthreads)
    {
      int t = 0;
#ifdef _OPENMP
      t = omp_get_thread_num();
#endif
      double* cp = priv + nF * (size_t)t;
      for (size_t q = 0; q < nF; q++) cp[q] = 0.0;
      int r0, r1;
      row_range(m, t, nthreads, &r0, &r1);
      for (int i = r0; i < r1; i++) {
        const double* er = E + (size_t)i * F;
        for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) {
          const double v = (double)val[k];
          double* cr = cp + (size_t)col[k] * F;
          for (int f = 0; f < F; f++) cr[f] += v * er[f];
        }
      }
#pragma omp barrier
#pragma omp for schedule(static)
      for (int j = 0; j < n; j++)
        for (int f = 0; f < F; f++) {
          double acc = 0.0;
          for (int q = 0; q < nthreads; q++) acc += priv[nF * (size_t)q + (size_t)j * F + f];
          C[(size_t)j * F + f] = acc;
        }
    }
    /* update x <- x c / (s + [s == 0]) and normalisation (reco.cu:273-281) */
    for (int f = 0; f < F; f++) {
      if (it >= niter_f[f]) continue;
      double sum = 0.0;
      for (int j = 0; j < n; j++) {
        const double dvs = s[j] + (s[j] == 0.0 ? 1.0 : 0.0);
        double* xj = &X[(size_t)j * F + f];
        *xj = *xj * C[(size_t)j * F + f] / dvs;
        sum += *xj;
      }
      for (int j = 0; j < n; j++) X[(size_t)j * F + f] /= sum;
    }
  }
  free(s); free(E); free(C); free(priv);
}

/* Back projection of an all-ones measurement in FP64 (tests/backprojection.cu:
 * 115-161 computes it in FP32 through the chunked SpMV^T). */
void eo_colsum_f64(int m, int n, const float* val, const int64_t* rowptr, const int* col,
                   double* s, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  const size_t nn = (size_t)(n > 0 ? n : 1);
  double* priv = (double*)calloc(nn * (size_t)nthreads, sizeof(double));
#pragma omp parallel num_threads(nthreads)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    double* cp = priv + nn * (size_t)t;
    int r0, r1;
    row_range(m, t, nthreads, &r0, &r1);
    for (int i = r0; i < r1; i++)
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) cp[col[k]] += (double)val[k];
#pragma omp barrier
#pragma omp for schedule(static)
    for (int j = 0; j < n; j++) {
      double acc = 0.0;
      for (int q = 0; q < nthreads; q++) acc += priv[nn * (size_t)q + j];
      s[j] = acc;
    }
  }
  free(priv);
}

void eo_mlem_iter_omp(int m, int n, const float* val, const int64_t* rowptr,
                      const int* col, const float* y, const float* s, float* x,
                      float* yhat, float* c, int nthreads) {
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
  nthreads = 1;
#endif
  /* forward projection and ratio, parallel over rows */
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (int i = 0; i < m; i++) {
    float acc = 0.0f;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) acc += val[k] * x[col[k]];
    yhat[i] = acc > 0.0f ? y[i] / acc : 0.0f;
  }
  /* back projection: per-thread private images, then a reduction */
  float* priv = (float*)calloc((size_t)nthreads * (size_t)n, sizeof(float));
#pragma omp parallel num_threads(nthreads)
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    float* cp = priv + (size_t)t * (size_t)n;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < m; i++) {
      float e = yhat[i];
      for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) cp[col[k]] += val[k] * e;
    }
#pragma omp for schedule(static)
    for (int j = 0; j < n; j++) {
      float acc = 0.0f;
      for (int q = 0; q < nthreads; q++) acc += priv[(size_t)q * (size_t)n + j];
      c[j] = acc;
    }
  }
  free(priv);
  double sum = 0.0;
  for (int j = 0; j < n; j++) {
    float dvs = s[j];
    dvs += (float)(dvs == 0.0f);
    x[j] = x[j] * c[j] / dvs;
    sum += x[j];
  }
  float sc = (float)(1.0 / sum);
  for (int j = 0; j < n; j++) x[j] *= sc;
}

double eo_loglik(int m, const float* val, const int64_t* rowptr, const int* col,
                 const float* y, const double* x) {
  double L = 0.0;
  for (int i = 0; i < m; i++) {
    double acc = 0.0;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; k++) acc += (double)val[k] * x[col[k]];
    if (acc > 0.0) L += (double)y[i] * log(acc) - acc;
  }
  return L;
}

/* ------------------------------------------------------------------------ */
/* chunk helpers: wrappers.hpp:405-445                                      */
/* ------------------------------------------------------------------------ */
uint32_t eo_nInChunk(uint32_t idChunk, uint32_t nAll, uint32_t nFullChunk) {
  if (idChunk < (nAll + nFullChunk - 1) / nFullChunk) {
    if (idChunk < nAll / nFullChunk) return nFullChunk;
    return nAll % nFullChunk;
  }
  return 0;
}
uint32_t eo_chunkPtr(uint32_t chunkId, uint32_t nFullChunk) { return chunkId * nFullChunk; }
uint32_t eo_nChunks(uint64_t maxNnz, uint64_t maxNChunk) {
  return (uint32_t)((maxNnz + maxNChunk - 1) / maxNChunk);
}

int eo_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
