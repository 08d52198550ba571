// Drop-in C++ surface test, written like the reference's test programs.
//   test_dropin          host-only checks (geometry functors, list, chunk helpers)
//   test_dropin --gpu    adds the device calls: getWorkqueue with capped and
//                        resumed cursors (test_getWorkqueue.cu:104-198), and the
//                        chunked MLEM loop of tests/reco.cu:209-300 through
//                        systemMatrixCalculation + CSRmv + mlemOperations.
// Prints one line per check and returns 0 when all pass.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "epetreco/epetreco.hpp"

static int g_fail = 0;
#define CHECK(cond, what)                                             \
  do {                                                                \
    if (cond) std::printf("ok   %s\n", what);                         \
    else { std::printf("FAIL %s (%s:%d)\n", what, __FILE__, __LINE__); g_fail++; } \
  } while (0)

// small workload: 5x5 segments per head, 18 angles of 10 deg, 16^3 grid of 3 mm
static MS small_setup() { return MS(-0.230f, 0.230f, 18, 5, 5, 5, 5, 10.f, 0.020f, 0.004f, 0.004f); }
static VG small_grid() { return VG(-0.024f, -0.024f, -0.024f, 0.003f, 0.003f, 0.003f, 16, 16, 16); }

static void host_checks() {
  MS setup = small_setup();
  VG grid = small_grid();
  const int nch = 18 * 5 * 5 * 5 * 5;
  bool rt = true;
  for (int lin = 0; lin < nch; lin++) {   // test_MeasurementSetupLinIndex.cu:37-46
    const int id = DefaultMeasurementSetupLinId<MS>()(Id0z()(lin, &setup), Id0y()(lin, &setup),
                                                      Id1z()(lin, &setup), Id1y()(lin, &setup),
                                                      Ida()(lin, &setup), &setup);
    rt = rt && id == lin;
  }
  CHECK(rt, "channel id round trip");
  rt = true;
  for (int lin = 0; lin < 16 * 16 * 16; lin++)   // test_VoxelGridLinIndex.cu:37-44
    rt = rt && DefaultVoxelGridLinId<VG>()(Idx()(lin, &grid), Idy()(lin, &grid),
                                           Idz()(lin, &grid), &grid) == lin;
  CHECK(rt, "voxel id round trip (z fastest)");
  // pixel centre of detector 0 at angle 0 is (pos0x, y, z) of the segment
  float c[3], b[3] = {0.5f, 0.5f, 0.5f};
  Trafo0()(c, b, 2, 2, 0, &setup);
  CHECK(std::fabs(c[0] + 0.230f) < 1e-7f && std::fabs(c[1]) < 1e-9f && std::fabs(c[2]) < 1e-9f,
        "trafo of the central pixel centre");
  float ci[3] = {0.5f, 0.5f, 0.5f};
  Trafo0_inplace()(ci, 2, 2, 0, &setup);
  CHECK(ci[0] == c[0] && ci[1] == c[1] && ci[2] == c[2], "inplace trafo equals trafo");
  ML list(3);
  list.set(0, 7, 1.f);
  bool threw = false;
  try { list.set(3, 1, 1.f); } catch (const std::out_of_range&) { threw = true; }
  CHECK(threw && list.cnlId(0) == 7, "MeasurementList::set bounds");
  CHECK((nInChunk<ListSizeType, ChunkGridSizeType>(3u, 1000u, 300u) == 100u) &&
            (nChunks<ChunkGridSizeType, MemArrSizeType>(1000ull, 300ull) == 4u) &&
            (chunkPtr<ListSizeType, ChunkGridSizeType>(2u, 300u) == 600u),
        "chunk helpers");
  float a[3] = {0, 0, 0}, bb[3] = {1, 0, 0}, p[3] = {0.5f, 2.f, 0};
  CHECK(distance(a, bb, p) == 2.f, "distance point-line");
}

static void gpu_checks() {
  MS setup = small_setup();
  VG grid = small_grid();
  const int nch = 18 * 5 * 5 * 5 * 5;
  const int V = 16 * 16 * 16;
  ML list(200);
  for (int i = 0; i < 200; i++) list.set(i, 37 * i, 1.f);
  // test_getWorkqueue.cu test3: a capped run is a prefix of the full run
  std::vector<int> c1, v1, c2, v2;
  int l1 = 0, x1 = 0, l2 = 0, x2 = 0;
  const int n1 = getWorkqueue<val_t, ML, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y, Ida,
                              Trafo0, Trafo1>(c1, v1, l1, x1, &list, &grid, &setup);
  const int n2 = getWorkqueue<val_t, ML, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y, Ida,
                              Trafo0, Trafo1>(c2, v2, l2, x2, &list, &grid, &setup, 5000);
  bool prefix = n2 == 5000 && n1 > n2;
  for (int i = 0; prefix && i < n2; i++) prefix = c1[i] == c2[i] && v1[i] == v2[i];
  CHECK(prefix, "capped workqueue is a prefix of the full one");
  CHECK(l1 == 200 && x1 == 0, "full scan leaves the cursor at the list end");
  // resume from the cursor until the end: concatenation == full run
  std::vector<int> cc = c2, vv = v2;
  while (l2 < 200) {
    std::vector<int> c3, v3;
    getWorkqueue<val_t, ML, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y, Ida, Trafo0, Trafo1>(
        c3, v3, l2, x2, &list, &grid, &setup, 7777);
    cc.insert(cc.end(), c3.begin(), c3.end());
    vv.insert(vv.end(), v3.begin(), v3.end());
  }
  CHECK(cc == c1 && vv == v1, "resumed workqueue == full workqueue");
  bool threw = false;
  try {
    getWorkqueue<val_t, ML, VG, Idx, Idy, Idz, MS, Id0z, Id0y, Id1z, Id1y, Ida, Trafo0, Trafo1>(
        c1, v1, l1, x1, &list, &grid, &setup);
  } catch (const epetreco::Error&) { threw = true; }
  CHECK(threw, "non-empty output vectors rejected");

  // ---- chunked MLEM (tests/reco.cu:209-300 semantics) on a synthetic y
  epetreco::Handle& handle = epetreco::defaultHandle();
  handle.setGeometry(setup, grid, 8, 0);
  const ListSizeType effM = 600;
  std::vector<int> yRowId_host(effM);
  std::vector<val_t> yVal_host(effM);
  for (ListSizeType i = 0; i < effM; i++) { yRowId_host[i] = 13 * i + 5; yVal_host[i] = 1.f + (i % 7); }
  const ListSizeType LIMM = 256;   // channels per chunk
  int* yRowId_devi; val_t* yVal_devi;
  epetreco::mallocD(yRowId_devi, effM); epetreco::mallocD(yVal_devi, effM);
  epetreco::memcpyH2D(yRowId_devi, yRowId_host.data(), effM);
  epetreco::memcpyH2D(yVal_devi, yVal_host.data(), effM);
  int *aCnlId, *aCsrCnlPtr, *aEcsrCnlPtr, *aVxlId; val_t* aVal; MemArrSizeType* nnz_devi;
  epetreco::mallocD(aCnlId, (size_t)LIMM * V); epetreco::mallocD(aCsrCnlPtr, nch + 1);
  epetreco::mallocD(aEcsrCnlPtr, LIMM + 1); epetreco::mallocD(aVxlId, (size_t)LIMM * V);
  epetreco::mallocD(aVal, (size_t)LIMM * V); epetreco::mallocD(nnz_devi, 1);
  val_t *x_devi, *xx_devi, *c_devi, *s_devi, *yT, *e;
  epetreco::mallocD(x_devi, V); epetreco::mallocD(xx_devi, V); epetreco::mallocD(c_devi, V);
  epetreco::mallocD(s_devi, V); epetreco::mallocD(yT, LIMM); epetreco::mallocD(e, LIMM);
  std::vector<val_t> ones(V, 1.f);
  epetreco::memcpyH2D(x_devi, ones.data(), V);
  scales<val_t>(x_devi, val_t(1. / sum<val_t>(x_devi, V)), V);
  epetreco::memcpyH2D(s_devi, ones.data(), V);
  scales<val_t>(s_devi, val_t(1. / sum<val_t>(s_devi, V)), V);
  epetreco::MatDescr A;
  val_t zero = 0.f, one = 1.f;
  const ChunkGridSizeType NChunks =
      nChunks<ChunkGridSizeType, MemArrSizeType>(MemArrSizeType(effM) * V, MemArrSizeType(LIMM) * V);
  std::vector<val_t> zeros(V, 0.f);
  for (int it = 0; it < 3; it++) {
    epetreco::memcpyH2D(c_devi, zeros.data(), V);
    for (ChunkGridSizeType chunk = 0; chunk < NChunks; chunk++) {
      ListSizeType m = nInChunk(chunk, effM, LIMM);
      ListSizeType ptr = chunkPtr(chunk, LIMM);
      MemArrSizeType nnz_host[1] = {0};
      epetreco::memcpyH2D(nnz_devi, nnz_host, 1);
      systemMatrixCalculation<val_t, ListSizeType, int, MemArrSizeType>(
          aEcsrCnlPtr, aVxlId, aVal, nnz_devi, aCnlId, aCsrCnlPtr, &(yRowId_devi[ptr]), &m,
          handle, nch, (long long)LIMM * V);
      epetreco::memcpyD2H(nnz_host, nnz_devi, 1);
      CSRmv<val_t>()(handle, epetreco::Operation::NON_TRANSPOSE, m, V, (int)*nnz_host, &one, A,
                     aVal, aEcsrCnlPtr, aVxlId, x_devi, &zero, yT);
      divides<val_t>(e, &(yVal_devi[ptr]), yT, m);
      CSRmv<val_t>()(handle, epetreco::Operation::TRANSPOSE, m, V, (int)*nnz_host, &one, A,
                     aVal, aEcsrCnlPtr, aVxlId, e, &one, c_devi);
    }
    dividesMultiplies<val_t>(xx_devi, x_devi, c_devi, s_devi, V);
    epetreco::memcpyD2D(x_devi, xx_devi, V);
    scales<val_t>(x_devi, val_t(1. / sum<val_t>(x_devi, V)), V);
  }
  std::vector<val_t> x(V);
  epetreco::memcpyD2H(x.data(), x_devi, V);
  double s = 0, mx = 0;
  bool finite = true;
  for (float v : x) { s += v; mx = std::max(mx, (double)v); finite = finite && std::isfinite(v) && v >= 0; }
  CHECK(finite && std::fabs(s - 1.0) < 1e-5 && mx > 1.0 / V, "chunked MLEM loop runs, sum x = 1");
  // write the image for the Python side to compare with the oracle
  FILE* f = std::fopen("dropin_mlem_x.bin", "wb");
  if (f) { std::fwrite(x.data(), sizeof(float), V, f); std::fclose(f); }
}

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "--gpu") == 0;
  try {
    host_checks();
    if (gpu) gpu_checks();
  } catch (const std::exception& ex) {
    std::printf("FAIL exception: %s\n", ex.what());
    return 2;
  }
  std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "PASSED", g_fail);
  return g_fail ? 1 : 0;
}
