// geometry.hpp -- value types and index/transform functors of the epetreco
// host surface (drop-in names and signatures, host code).
//
//   DefaultVoxelGrid<T>            VoxelGrid.hpp:385-497
//   DefaultVoxelGridLinId/Idx/...  VoxelGridLinIndex.hpp:193-282 (z fastest)
//   DefaultMeasurementSetup<T>     MeasurementSetup.hpp:124-294
//   DefaultMeasurementSetupLinId.. MeasurementSetupLinIndex.hpp:32-165
//   ...Trafo2CartCoord*            MeasurementSetupTrafo2CartCoord.hpp:42-360
//   DefaultMeasurementList<T>      MeasurementList.hpp:29-80
//   absolute/crossProduct/distance distancePointLine.h:19-55
//
// Member layouts equal the reference's, so a DefaultMeasurementSetup<float> /
// DefaultVoxelGrid<float> is bit-compatible with epet_setup / epet_grid.  The
// functors evaluate the reference's single-precision expressions in the
// reference's order (host float arithmetic).
#pragma once

#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../epet.h"

namespace epetreco_detail {
inline constexpr double kPi = 3.14159265358979323846;
}

template <typename T>
class DefaultVoxelGrid {
 public:
  DefaultVoxelGrid() = default;
  DefaultVoxelGrid(T const gridO0, T const gridO1, T const gridO2, T const gridD0, T const gridD1,
                   T const gridD2, int const gridN0, int const gridN1, int const gridN2)
      : gridO{gridO0, gridO1, gridO2}, gridD{gridD0, gridD1, gridD2},
        gridN{gridN0, gridN1, gridN2} {}
  T gridox() const { return gridO[0]; }
  T gridoy() const { return gridO[1]; }
  T gridoz() const { return gridO[2]; }
  T griddx() const { return gridD[0]; }
  T griddy() const { return gridD[1]; }
  T griddz() const { return gridD[2]; }
  int gridnx() const { return gridN[0]; }
  int gridny() const { return gridN[1]; }
  int gridnz() const { return gridN[2]; }
  // Deprecated in the reference (x fastest, VoxelGrid.hpp:477-496); not used
  // by the reconstruction path.
  int linVoxelId(int const* const s) const { return s[0] + s[1] * gridN[0] + s[2] * gridN[0] * gridN[1]; }
  void sepVoxelId(int* const s, int const lin) const {
    int t = lin;
    s[2] = t / (gridN[0] * gridN[1]);
    t %= (gridN[0] * gridN[1]);
    s[1] = t / gridN[0];
    s[0] = t % gridN[0];
  }
  epet_grid c_grid() const {
    epet_grid g;
    for (int d = 0; d < 3; d++) { g.o[d] = (float)gridO[d]; g.d[d] = (float)gridD[d]; g.n[d] = gridN[d]; }
    return g;
  }

 private:
  T gridO[3];
  T gridD[3];
  int gridN[3];
};

template <typename VG>
struct DefaultVoxelGridLinId {
  int operator()(int const idx, int const idy, int const idz, VG const* const vg) const {
    return idx * (vg->gridnz() * vg->gridny()) + idy * (vg->gridnz()) + idz;
  }
};
template <typename VG>
struct DefaultVoxelGridIdx {
  int operator()(int const linId, VG const* const vg) const {
    return linId / (vg->gridnz() * vg->gridny());
  }
};
template <typename VG>
struct DefaultVoxelGridIdy {
  int operator()(int const linId, VG const* const vg) const {
    return (linId % (vg->gridnz() * vg->gridny())) / vg->gridnz();
  }
};
template <typename VG>
struct DefaultVoxelGridIdz {
  int operator()(int const linId, VG const* const vg) const {
    return (linId % (vg->gridnz() * vg->gridny())) % vg->gridnz();
  }
};

template <typename T>
class DefaultMeasurementSetup {
 public:
  DefaultMeasurementSetup() = default;
  DefaultMeasurementSetup(T pos0x, T pos1x, int na, int n0z, int n0y, int n1z, int n1y, T da,
                          T segx, T segy, T segz)
      : _pos0x(pos0x), _pos1x(pos1x), _na(na), _da(da), _n0z(n0z), _n0y(n0y), _n1z(n1z),
        _n1y(n1y), _segx(segx), _segy(segy), _segz(segz) {}
  T pos0x() const { return _pos0x; }
  T pos1x() const { return _pos1x; }
  int na() const { return _na; }
  T da() const { return _da; }
  int n0z() const { return _n0z; }
  int n0y() const { return _n0y; }
  int n1z() const { return _n1z; }
  int n1y() const { return _n1y; }
  T segx() const { return _segx; }
  T segy() const { return _segy; }
  T segz() const { return _segz; }
  epet_setup c_setup() const {
    return epet_setup{(float)_pos0x, (float)_pos1x, _na, (float)_da, _n0z, _n0y, _n1z, _n1y,
                      (float)_segx, (float)_segy, (float)_segz};
  }

 private:
  T _pos0x, _pos1x;
  int _na;
  T _da;
  int _n0z, _n0y, _n1z, _n1y;
  T _segx, _segy, _segz;
};

template <typename MS>
struct DefaultMeasurementSetupLinId {
  int operator()(int const id0z, int const id0y, int const id1z, int const id1y, int const ida,
                 MS const* const m) const {
    return ida * (m->n1y() * m->n1z() * m->n0y() * m->n0z()) +
           id0z * (m->n1y() * m->n1z() * m->n0y()) + id0y * (m->n1y() * m->n1z()) +
           id1z * (m->n1y()) + id1y;
  }
};
template <typename MS>
struct DefaultMeasurementSetupIda {
  int operator()(int const lin, MS const* const m) const {
    return lin / (m->n1y() * m->n1z() * m->n0y() * m->n0z());
  }
};
template <typename MS>
struct DefaultMeasurementSetupId0z {
  int operator()(int const lin, MS const* const m) const {
    return (lin % (m->n1y() * m->n1z() * m->n0y() * m->n0z())) / (m->n1y() * m->n1z() * m->n0y());
  }
};
template <typename MS>
struct DefaultMeasurementSetupId0y {
  int operator()(int const lin, MS const* const m) const {
    return (lin % (m->n1y() * m->n1z() * m->n0y())) / (m->n1y() * m->n1z());
  }
};
template <typename MS>
struct DefaultMeasurementSetupId1z {
  int operator()(int const lin, MS const* const m) const {
    return (lin % (m->n1y() * m->n1z())) / m->n1y();
  }
};
template <typename MS>
struct DefaultMeasurementSetupId1y {
  int operator()(int const lin, MS const* const m) const { return lin % m->n1y(); }
};

namespace epetreco_detail {
// R_y(theta) (pos + (b - 1/2) e) with the reference's float order
// (MeasurementSetupTrafo2CartCoord.hpp:56-104).
template <typename T, typename MS>
void segment_trafo(T* const out, T const* const b, int const idz, int const idy, int const ida,
                   MS const* const m, int const det) {
  const T e0 = m->segx(), e1 = m->segy(), e2 = m->segz();
  const int ny = det ? m->n1y() : m->n0y();
  const int nz = det ? m->n1z() : m->n0z();
  const T p0 = det ? m->pos1x() : m->pos0x();
  const T p1 = (idy - T(.5) * (ny - 1)) * e1;
  const T p2 = (idz - T(.5) * (nz - 1)) * e2;
  const T sn = sinf(T(kPi) * ida * (m->da()) / T(180.0));
  const T cs = cosf(T(kPi) * ida * (m->da()) / T(180.0));
  const T t[12] = {cs * e0, T(0.), sn * e2, cs * (p0 - T(.5) * e0) + sn * (p2 - T(.5) * e2),
                   T(0.), e1, T(0.), p1 - T(.5) * e1,
                   -sn * e0, T(0.), cs * e2, -sn * (p0 - T(.5) * e0) + cs * (p2 - T(.5) * e2)};
  const T b0 = b[0], b1 = b[1], b2 = b[2];
  for (int r = 0; r < 3; r++)
    out[r] = t[4 * r + 0] * b0 + t[4 * r + 1] * b1 + t[4 * r + 2] * b2 + t[4 * r + 3] * T(1.);
}
}  // namespace epetreco_detail

template <typename T, typename MS>
struct DefaultMeasurementSetupTrafo2CartCoordFirstPixel {
  void operator()(T* const cart, T const* const box, int const id0z, int const id0y,
                  int const ida, MS const* const m) const {
    epetreco_detail::segment_trafo(cart, box, id0z, id0y, ida, m, 0);
  }
};
template <typename T, typename MS>
struct DefaultMeasurementSetupTrafo2CartCoordSecndPixel {
  void operator()(T* const cart, T const* const box, int const id1z, int const id1y,
                  int const ida, MS const* const m) const {
    epetreco_detail::segment_trafo(cart, box, id1z, id1y, ida, m, 1);
  }
};
template <typename T, typename MS>
struct DefaultMeasurementSetupTrafo2CartCoordFirstPixel_inplace {
  void operator()(T* const c, int const id0z, int const id0y, int const ida,
                  MS const* const m) const {
    const T tmp[3] = {c[0], c[1], c[2]};
    epetreco_detail::segment_trafo(c, tmp, id0z, id0y, ida, m, 0);
  }
};
template <typename T, typename MS>
struct DefaultMeasurementSetupTrafo2CartCoordSecndPixel_inplace {
  void operator()(T* const c, int const id1z, int const id1y, int const ida,
                  MS const* const m) const {
    const T tmp[3] = {c[0], c[1], c[2]};
    epetreco_detail::segment_trafo(c, tmp, id1z, id1y, ida, m, 1);
  }
};

// DefaultMeasurementList (MeasurementList.hpp:29-80).  An out-of-range set()
// throws std::out_of_range where the reference exits.
template <typename T>
class DefaultMeasurementList {
 public:
  explicit DefaultMeasurementList(int const size) : _cnlId(size), _val(size), _size(size) {}
  int cnlId(int const listId) const { return _cnlId[listId]; }
  T val(int const listId) const { return _val[listId]; }
  int size() const { return _size; }
  void set(int const listId, int const cnlId, T const val) {
    if (listId >= _size || listId < 0)
      throw std::out_of_range("DefaultMeasurementList::set(...) : error : listId out of bounds");
    _cnlId[listId] = cnlId;
    _val[listId] = val;
  }
  const int* cnlIds() const { return _cnlId.data(); }

 private:
  std::vector<int> _cnlId;
  std::vector<T> _val;
  int _size;
};

// distancePointLine.h:19-55
template <typename T>
inline T absolute(T const* const a) { return sqrtf(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
template <typename T>
inline T scalarProduct(T const* const a, T const* const b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
template <typename T>
inline void crossProduct(T* const r, T const* const a, T const* const b) {
  for (int d = 0; d < 3; d++) r[d] = a[(d + 1) % 3] * b[(d + 2) % 3] - a[(d + 2) % 3] * b[(d + 1) % 3];
}
template <typename T>
T distance(T const* const a, T const* const b, T const* const p) {
  T ap[3], ab[3], c[3];
  for (int d = 0; d < 3; d++) { ap[d] = p[d] - a[d]; ab[d] = b[d] - a[d]; }
  crossProduct(c, ab, ap);
  return absolute(c) / absolute(ab);
}
