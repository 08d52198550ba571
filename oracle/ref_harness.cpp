// ref_harness.cpp -- C entry points around the reference's own host-side
// geometry and workqueue templates, compiled from the reference sources where
// they lie (include path set by oracle/Makefile) into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: the checker for the oracle's restatement.  It is
// built with hipcc in host-only mode, so the reference's __host__/__device__
// qualifiers come from the installed HIP headers; nothing is stubbed.  Only
// header-only, CUDA-free parts of the reference are used:
//   VoxelGrid.hpp, VoxelGridLinIndex.hpp, MeasurementSetup.hpp,
//   MeasurementSetupLinIndex.hpp, MeasurementSetupTrafo2CartCoord.hpp,
//   MeasurementList.hpp, distancePointLine.h, getWorkqueue.hpp,
//   typedefs_geometry.hpp.
#include <vector>
#include <cstring>

#include "typedefs_geometry.hpp"
#include "getWorkqueue.hpp"

#include "epet_oracle.h"   // eo_setup / eo_grid field layout only

namespace {
MS make_setup(const eo_setup* s) {
  return MS(s->pos0x, s->pos1x, s->na, s->n0z, s->n0y, s->n1z, s->n1y,
            s->da, s->segx, s->segy, s->segz);
}
VG make_grid(const eo_grid* g) {
  return VG(g->o[0], g->o[1], g->o[2], g->d[0], g->d[1], g->d[2],
            g->n[0], g->n[1], g->n[2]);
}
}  // namespace

extern "C" {

// getWorkqueue (getWorkqueue.hpp:129-171 with cap >= 0, :187-223 with cap < 0).
// Copies at most out_cap pairs; returns the reference's nFound.
long long ref_workqueue(const eo_setup* s, const eo_grid* g, const int* list,
                        int list_size, int* listId, int* vxlId, int cap,
                        int* out_cnl, int* out_vxl, long long out_cap) {
  MS setup = make_setup(s);
  VG grid = make_grid(g);
  ML ml(list_size);
  for (int i = 0; i < list_size; i++) ml.set(i, list[i], val_t(1));
  std::vector<int> cnl, vxl;
  int n;
  if (cap >= 0)
    n = getWorkqueue<val_t, ML, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y,
                     Ida, Trafo0, Trafo1>(cnl, vxl, *listId, *vxlId, &ml, &grid,
                                          &setup, cap);
  else
    n = getWorkqueue<val_t, ML, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y,
                     Ida, Trafo0, Trafo1>(cnl, vxl, *listId, *vxlId, &ml, &grid,
                                          &setup);
  long long k = (long long)cnl.size() < out_cap ? (long long)cnl.size() : out_cap;
  if (out_cnl && k > 0) std::memcpy(out_cnl, cnl.data(), sizeof(int) * k);
  if (out_vxl && k > 0) std::memcpy(out_vxl, vxl.data(), sizeof(int) * k);
  return n;
}

// Segment transform, non-inplace (Trafo0/Trafo1) or inplace variants.
void ref_trafo(const eo_setup* s, int det, int inplace, const float box[3],
               int idz, int idy, int ida, float out[3]) {
  MS setup = make_setup(s);
  if (!inplace) {
    if (det == 0) Trafo0()(out, box, idz, idy, ida, &setup);
    else          Trafo1()(out, box, idz, idy, ida, &setup);
  } else {
    out[0] = box[0]; out[1] = box[1]; out[2] = box[2];
    if (det == 0) Trafo0_inplace()(out, idz, idy, ida, &setup);
    else          Trafo1_inplace()(out, idz, idy, ida, &setup);
  }
}

float ref_distance(const float a[3], const float b[3], const float p[3]) {
  return distance<float>(a, b, p);
}

// Channel id <-> (ida, id0z, id0y, id1z, id1y) functors.
int ref_ms_linid(const eo_setup* s, int id0z, int id0y, int id1z, int id1y, int ida) {
  MS setup = make_setup(s);
  return DefaultMeasurementSetupLinId<MS>()(id0z, id0y, id1z, id1y, ida, &setup);
}
void ref_ms_decompose(const eo_setup* s, int lin, int out[5]) {
  MS setup = make_setup(s);
  out[0] = Ida()(lin, &setup);
  out[1] = Id0z()(lin, &setup);
  out[2] = Id0y()(lin, &setup);
  out[3] = Id1z()(lin, &setup);
  out[4] = Id1y()(lin, &setup);
}

// Voxel id <-> (idx, idy, idz) functors.
int ref_vg_linid(const eo_grid* g, int idx, int idy, int idz) {
  VG grid = make_grid(g);
  return DefaultVoxelGridLinId<VG>()(idx, idy, idz, &grid);
}
void ref_vg_decompose(const eo_grid* g, int lin, int out[3]) {
  VG grid = make_grid(g);
  out[0] = Idx()(lin, &grid);
  out[1] = Idy()(lin, &grid);
  out[2] = Idz()(lin, &grid);
}

}  // extern "C"
