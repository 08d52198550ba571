// workloads.hpp -- scanner and grid of the named workloads (the same values as
// epetreco_amd/geometry.py): two heads 460 mm apart, segments 20 x 4 x 4 mm.
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>

#include "epet.h"

namespace epet_drv {

inline void workload(const std::string& name, epet_setup* s, epet_grid* g) {
  std::memset(s, 0, sizeof(*s));
  std::memset(g, 0, sizeof(*g));
  s->pos0x = -0.230f;
  s->pos1x = 0.230f;
  s->segx = 0.020f;
  s->segy = 0.004f;
  s->segz = 0.004f;
  auto grid = [g](float o, float d, int n) {
    for (int k = 0; k < 3; k++) {
      g->o[k] = o;
      g->d[k] = d;
      g->n[k] = n;
    }
  };
  if (name == "small" || name == "projector") {
    s->na = 18;
    s->da = 10.0f;
    s->n0z = s->n0y = s->n1z = s->n1y = 5;
    if (name == "small") grid(-0.024f, 0.003f, 16);
    else grid(-0.026f, 0.001f, 52);
  } else if (name == "full" || name == "large") {
    s->na = 180;
    s->da = 2.0f;
    s->n0z = s->n0y = s->n1z = s->n1y = 13;
    if (name == "full") grid(-0.026f, 0.001f, 52);
    else grid(-0.026f, 0.0005f, 104);
  } else {
    throw std::invalid_argument("unknown workload " + name);
  }
}

inline long long nchannels(const epet_setup& s) {
  return (long long)s.na * s.n0z * s.n0y * s.n1z * s.n1y;
}

}  // namespace epet_drv
