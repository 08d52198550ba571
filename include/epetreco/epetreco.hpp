// epetreco.hpp -- C++ host surface of the MLEM hot path over the C-ABI
// (include/epet.h, libepet.so).  Names, template parameters and argument
// orders follow the reference's headers so that a caller of the reference can
// switch by changing includes:
//
//   getWorkqueue / getWorkqueueEntry      getWorkqueue.hpp:33-37, :129-135, :187-192
//   getSystemMatrixFromWorkqueue          getSystemMatrixFromWorkqueue.cu:97-102
//   getSystemMatrix                       getSystemMatrixDeviceOnly.cu:324-331
//   systemMatrixCalculation               wrappers.hpp:172-177
//   cooSort                               cooSort.hpp:53-59
//   convertCoo2Csr, customizeMatDescr     cusparseWrapper.hpp:21-26, :38-42
//   convertCsr2Ecsr                       convertCsr2Ecsr.hpp:34-39
//   CSRmv<float>                          csrmv.hpp:15-35
//   divides, dividesMultiplies, scales, sum   mlemOperations.hpp:22-139
//   nInChunk, chunkPtr, nChunks           wrappers.hpp:405-445
//   mallocD, memcpyH2D/D2H/D2D            cuda_wrappers.hpp:15-87
//
// Device errors throw epetreco::Error with the library's file(line) message,
// keeping HANDLE_ERROR's throwing contract (CUDA_HandleError.hpp:17-31).  The
// cuSPARSE handle and matrix descriptor become epetreco::Handle (one context:
// device + stream + geometry) and epetreco::MatDescr (general, zero-based).
// The reference's compile-time geometry in __constant__ memory becomes
// Handle::setGeometry (device_constant_memory.hpp:14-26, tests/reco.cu:64-72).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../epet.h"
#include "geometry.hpp"

namespace epetreco {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc) {
  if (rc != EPET_OK) throw Error(rc, epet_last_error());
}

// Replaces cusparseHandle_t plus the implicit CUDA device/stream.
class Handle {
 public:
  explicit Handle(int device = 0, void* stream = nullptr) {
    check(epet_ctx_create(&ctx_, device, stream));
  }
  ~Handle() { epet_ctx_destroy(ctx_); }
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  epet_ctx* get() const { return ctx_; }
  template <typename MS, typename VG>
  void setGeometry(MS const& setup, VG const& grid, int nrays, uint64_t seed = 0) {
    const epet_setup s = setup.c_setup();
    const epet_grid g = grid.c_grid();
    check(epet_set_geometry(ctx_, &s, &g));
    check(epet_set_nrays(ctx_, nrays));
    check(epet_set_seed(ctx_, seed));
  }
  void synchronize() const { check(epet_ctx_synchronize(ctx_)); }

 private:
  epet_ctx* ctx_ = nullptr;
};

// Process-wide default handle, as the reference relies on the default device.
inline Handle& defaultHandle() {
  static Handle h(0, nullptr);
  return h;
}

enum class Operation { NON_TRANSPOSE = EPET_OP_N, TRANSPOSE = EPET_OP_T };
struct MatDescr {};   // general matrix, zero-based indices (cusparseWrapper.hpp:21-26)
inline int customizeMatDescr(MatDescr&, Handle const&) { return EPET_OK; }

// ---- memory (cuda_wrappers.hpp:15-87) -----------------------------------
template <typename T>
void mallocD(T*& devi, size_t n, Handle& h = defaultHandle()) {
  void* p = nullptr;
  check(epet_malloc(h.get(), &p, sizeof(T) * n));
  devi = static_cast<T*>(p);
}
template <typename T>
void freeD(T* devi, Handle& h = defaultHandle()) { check(epet_free(h.get(), devi)); }
template <typename T>
void memcpyH2D(T* devi, T const* host, size_t n, Handle& h = defaultHandle()) {
  check(epet_memcpy_h2d(h.get(), devi, host, sizeof(T) * n));
}
template <typename T>
void memcpyD2H(T* host, T const* devi, size_t n, Handle& h = defaultHandle()) {
  check(epet_memcpy_d2h(h.get(), host, devi, sizeof(T) * n));
}
template <typename T>
void memcpyD2D(T* d0, T const* d1, size_t n, Handle& h = defaultHandle()) {
  check(epet_memcpy_d2d(h.get(), d0, d1, sizeof(T) * n));
}

}  // namespace epetreco

// ---- workqueue (host vectors, device scan) --------------------------------
namespace epetreco_detail {
template <typename ConcreteList>
inline int* upload_list(ConcreteList const* list, epet_ctx* ctx) {
  std::vector<int> ids(list->size());
  for (int i = 0; i < list->size(); i++) ids[i] = list->cnlId(i);
  void* d = nullptr;
  epetreco::check(epet_malloc(ctx, &d, sizeof(int) * (ids.size() ? ids.size() : 1)));
  if (!ids.empty()) epetreco::check(epet_memcpy_h2d(ctx, d, ids.data(), sizeof(int) * ids.size()));
  return static_cast<int*>(d);
}

template <typename ConcreteList, typename VG, typename MS>
int workqueue_impl(std::vector<int>& cnl, std::vector<int>& vxl, int& listId, int& vxlId,
                   ConcreteList const* const list, VG const* const grid, MS const* const setup,
                   long long cap) {
  if (!cnl.empty() || !vxl.empty())
    throw epetreco::Error(EPET_ERR_ARG, "getWorkqueue: error: return vector not empty!");
  epetreco::Handle& h = epetreco::defaultHandle();
  const epet_setup s = setup->c_setup();
  const epet_grid g = grid->c_grid();
  epetreco::check(epet_set_geometry(h.get(), &s, &g));
  int* dl = upload_list(list, h.get());
  const long long V = (long long)grid->gridnx() * grid->gridny() * grid->gridnz();
  (void)V;
  // size the output with a counting query (cursor unchanged)
  long long capacity = 0;
  {
    int li = listId, vi = vxlId;
    const int rc0 = epet_getWorkqueue(h.get(), dl, list->size(), &li, &vi, cap, nullptr, nullptr,
                                      0, &capacity);
    if (rc0 != EPET_OK) { epet_free(h.get(), dl); epetreco::check(rc0); }
  }
  void *dc = nullptr, *dv = nullptr;
  long long n = 0;
  const long long alloc = capacity > 0 ? capacity : 1;
  epetreco::check(epet_malloc(h.get(), &dc, sizeof(int) * alloc));
  epetreco::check(epet_malloc(h.get(), &dv, sizeof(int) * alloc));
  const int rc = epet_getWorkqueue(h.get(), dl, list->size(), &listId, &vxlId, cap,
                                   static_cast<int*>(dc), static_cast<int*>(dv), capacity, &n);
  if (rc == EPET_OK) {
    cnl.resize(n);
    vxl.resize(n);
    if (n) {
      epetreco::check(epet_memcpy_d2h(h.get(), cnl.data(), dc, sizeof(int) * n));
      epetreco::check(epet_memcpy_d2h(h.get(), vxl.data(), dv, sizeof(int) * n));
    }
  }
  epet_free(h.get(), dc);
  epet_free(h.get(), dv);
  epet_free(h.get(), dl);
  epetreco::check(rc);
  return (int)n;
}
}  // namespace epetreco_detail

// getWorkqueue.hpp:129-171 (capped at n)
template <typename T, typename ConcreteList, typename ConcreteVG, typename ConcreteVGidx,
          typename ConcreteVGidy, typename ConcreteVGidz, typename ConcreteMS,
          typename ConcreteMSid0z, typename ConcreteMSid0y, typename ConcreteMSid1z,
          typename ConcreteMSid1y, typename ConcreteMSida, typename ConcreteTrafo0,
          typename ConcreteTrafo1>
int getWorkqueue(std::vector<int>& returnCnlId, std::vector<int>& returnVxlId, int& listId,
                 int& vxlId, ConcreteList const* const list, ConcreteVG const* const grid,
                 ConcreteMS const* const setup, int const n) {
  return epetreco_detail::workqueue_impl(returnCnlId, returnVxlId, listId, vxlId, list, grid,
                                         setup, n < 0 ? 0 : n);
}

// getWorkqueue.hpp:187-223 (complete)
template <typename T, typename ConcreteList, typename ConcreteVG, typename ConcreteVGidx,
          typename ConcreteVGidy, typename ConcreteVGidz, typename ConcreteMS,
          typename ConcreteMSid0z, typename ConcreteMSid0y, typename ConcreteMSid1z,
          typename ConcreteMSid1y, typename ConcreteMSida, typename ConcreteTrafo0,
          typename ConcreteTrafo1>
int getWorkqueue(std::vector<int>& returnCnlId, std::vector<int>& returnVxlId, int& listId,
                 int& vxlId, ConcreteList const* const list, ConcreteVG const* const grid,
                 ConcreteMS const* const setup) {
  return epetreco_detail::workqueue_impl(returnCnlId, returnVxlId, listId, vxlId, list, grid,
                                         setup, -1);
}

// getWorkqueue.hpp:33-113: the next candidate from the cursor; returns 1/0.
template <typename T, typename ConcreteList, typename ConcreteVG, typename ConcreteVGidx,
          typename ConcreteVGidy, typename ConcreteVGidz, typename ConcreteMS,
          typename ConcreteMSid0z, typename ConcreteMSid0y, typename ConcreteMSid1z,
          typename ConcreteMSid1y, typename ConcreteMSida, typename ConcreteTrafo0,
          typename ConcreteTrafo1>
int getWorkqueueEntry(int* const returnCnlId, int* const returnVxlId, int& listId, int& vxlId,
                      ConcreteList const* const list, ConcreteVG const* const grid,
                      ConcreteMS const* const setup) {
  std::vector<int> c, v;
  const int n = epetreco_detail::workqueue_impl(c, v, listId, vxlId, list, grid, setup, 1);
  if (n == 1) { *returnCnlId = c[0]; *returnVxlId = v[0]; }
  return n;
}

// ---- system matrix ----------------------------------------------------------
// getSystemMatrixFromWorkqueue.cu:97-102 (device arrays; template parameters
// kept for signature compatibility, geometry comes from the handle).
template <typename T, typename... Functors>
void getSystemMatrixFromWorkqueue(int const* const globalCnl, int const* const globalVxl,
                                  T* const globalVal, int const wqLength, int const nrays,
                                  epetreco::Handle& h = epetreco::defaultHandle()) {
  static_assert(sizeof(T) == 4, "system-matrix values are FP32");
  epetreco::check(epet_getSystemMatrixFromWorkqueue(h.get(), globalCnl, globalVxl,
                                                    reinterpret_cast<float*>(globalVal),
                                                    wqLength, nrays));
}

// getSystemMatrixDeviceOnly.cu:324-331 as a host launcher: COO of every
// candidate of the listed channels (zeros included), appended at
// *truckDest_devi, row-major order.  capacity = length of the output arrays.
template <typename T, typename ListSizeType, typename GridSizeType, typename MemArrSizeType>
void getSystemMatrix(T* const sme_devi, GridSizeType* const vxlId_devi, int* const cnlId_devi,
                     int const* const ml_devi, ListSizeType const* const mlSize_devi,
                     MemArrSizeType* const truckDest_devi, long long capacity,
                     epetreco::Handle& h = epetreco::defaultHandle()) {
  static_assert(sizeof(GridSizeType) == 4 && sizeof(MemArrSizeType) == 8, "int32 ids, u64 count");
  ListSizeType mlSize = 0;
  epetreco::memcpyD2H(&mlSize, mlSize_devi, 1, h);
  epetreco::check(epet_getSystemMatrix(h.get(), reinterpret_cast<float*>(sme_devi),
                                       reinterpret_cast<int*>(vxlId_devi), cnlId_devi, ml_devi,
                                       (uint32_t)mlSize,
                                       reinterpret_cast<unsigned long long*>(truckDest_devi),
                                       capacity, 1));
}

// wrappers.hpp:172-247.  Buffer sizes as mallocD_SystemMatrix allocates them
// (wrappers.hpp:64-74): vxlId/val/cnlId of capacity nY * voxels, csrPtr of
// nchannels + 1, ecsrPtr of nY + 1.
template <typename T, typename ListSizeType, typename GridSizeType, typename MemArrSizeType,
          typename RayGen = void>
void systemMatrixCalculation(int* const aEcsrCnlPtr_devi, GridSizeType* const aVxlId_devi,
                             T* const aVal_devi, MemArrSizeType* const nnz_devi,
                             int* const aCnlId_devi, int* const aCsrCnlPtr_devi,
                             int* const yRowId_devi, ListSizeType const* const& nY_host,
                             epetreco::Handle& handle, int nchannels, long long capacity) {
  static_assert(sizeof(T) == 4 && sizeof(GridSizeType) == 4 && sizeof(MemArrSizeType) == 8,
                "FP32 values, int32 ids, u64 count");
  epetreco::check(epet_systemMatrixCalculation(
      handle.get(), aEcsrCnlPtr_devi, reinterpret_cast<int*>(aVxlId_devi),
      reinterpret_cast<float*>(aVal_devi), reinterpret_cast<unsigned long long*>(nnz_devi),
      aCnlId_devi, aCsrCnlPtr_devi, yRowId_devi, (uint32_t)nY_host[0], nchannels, capacity, 1));
}

// ---- sparse formats -----------------------------------------------------------
template <typename T>
void cooSort(T* const cooVal, int* const cooRowId, int* const cooColId, int const N,
             epetreco::Handle& h = epetreco::defaultHandle()) {
  static_assert(sizeof(T) == 4, "values are FP32");
  epetreco::check(epet_cooSort(h.get(), reinterpret_cast<float*>(cooVal), cooRowId, cooColId, N));
}

inline int convertCoo2Csr(int* const csrRowPtr, int const* const cooRowId,
                          epetreco::Handle const& handle, int const nnz, int const m) {
  return epet_convertCoo2Csr(handle.get(), csrRowPtr, cooRowId, nnz, m);
}

inline void convertCsr2Ecsr(int* const ecsrRowPtr, int* const yRowId, int const nnzr,
                            int* const csrRowPtr, int const m,
                            epetreco::Handle& h = epetreco::defaultHandle()) {
  epetreco::check(epet_convertCsr2Ecsr(h.get(), ecsrRowPtr, yRowId, nnzr, csrRowPtr, m));
}

// ---- CSRmv (csrmv.hpp:15-35) ---------------------------------------------------
template <typename T>
struct CSRmv;

template <>
struct CSRmv<float> {
  void operator()(epetreco::Handle const& handle, epetreco::Operation transA, int m, int n,
                  int nnz, float* const alpha, epetreco::MatDescr, float* const csrValA,
                  int* const csrRowPtrA, int* const csrColIdA, float* const x,
                  float* const beta, float* const y) const {
    epetreco::check(epet_csrmv(handle.get(), (int)transA, m, n, nnz, alpha, csrValA, csrRowPtrA,
                               csrColIdA, x, beta, y));
  }
};

// ---- mlemOperations.hpp:22-139 -------------------------------------------------
template <typename T>
void divides(T* const& result, T* const& dividend, T* const& divisor, int const n,
             epetreco::Handle& h = epetreco::defaultHandle()) {
  epetreco::check(epet_divides(h.get(), result, dividend, divisor, n));
}
template <typename T>
void dividesMultiplies(T* const& result, T* const& a, T* const& b, T* const& c, int const n,
                       epetreco::Handle& h = epetreco::defaultHandle()) {
  epetreco::check(epet_dividesMultiplies(h.get(), result, a, b, c, n));
}
template <typename T>
void scales(T* const& x, T const alpha, int const n,
            epetreco::Handle& h = epetreco::defaultHandle()) {
  epetreco::check(epet_scales(h.get(), x, alpha, n));
}
template <typename T>
T sum(T* const& x, int const n, epetreco::Handle& h = epetreco::defaultHandle()) {
  double s = 0.0;
  epetreco::check(epet_sum(h.get(), x, n, &s));
  return T(s);
}

// ---- chunk helpers (wrappers.hpp:405-445) ------------------------------------------
template <typename ListSizeType, typename ChunkGridSizeType>
ListSizeType nInChunk(ChunkGridSizeType idChunk, ListSizeType const nAll,
                      ListSizeType const nFullChunk) {
  return (ListSizeType)epet_nInChunk((uint32_t)idChunk, (uint32_t)nAll, (uint32_t)nFullChunk);
}
template <typename ListSizeType, typename ChunkGridSizeType>
ListSizeType chunkPtr(ChunkGridSizeType const chunkId, ListSizeType const nFullChunk) {
  return (ListSizeType)epet_chunkPtr((uint32_t)chunkId, (uint32_t)nFullChunk);
}
template <typename ChunkGridSizeType, typename MemArrSizeType>
ChunkGridSizeType nChunks(MemArrSizeType const maxNnz, MemArrSizeType const maxNChunk) {
  return (ChunkGridSizeType)epet_nChunks(maxNnz, maxNChunk);
}

// ---- type aliases (typedefs_geometry.hpp:23-119, typedefs_array_sizes.hpp) -----------
typedef float val_t;
typedef DefaultVoxelGrid<val_t> VG;
typedef DefaultVoxelGridIdx<VG> Idx;
typedef DefaultVoxelGridIdy<VG> Idy;
typedef DefaultVoxelGridIdz<VG> Idz;
typedef DefaultMeasurementSetup<val_t> MS;
typedef DefaultMeasurementSetupId0z<MS> Id0z;
typedef DefaultMeasurementSetupId0y<MS> Id0y;
typedef DefaultMeasurementSetupId1z<MS> Id1z;
typedef DefaultMeasurementSetupId1y<MS> Id1y;
typedef DefaultMeasurementSetupIda<MS> Ida;
typedef DefaultMeasurementList<val_t> ML;
typedef DefaultMeasurementSetupTrafo2CartCoordFirstPixel<val_t, MS> Trafo0;
typedef DefaultMeasurementSetupTrafo2CartCoordSecndPixel<val_t, MS> Trafo1;
typedef DefaultMeasurementSetupTrafo2CartCoordFirstPixel_inplace<val_t, MS> Trafo0_inplace;
typedef DefaultMeasurementSetupTrafo2CartCoordSecndPixel_inplace<val_t, MS> Trafo1_inplace;
typedef unsigned long long int MemArrSizeType;
typedef uint32_t ListSizeType;
typedef uint32_t ChunkGridSizeType;
