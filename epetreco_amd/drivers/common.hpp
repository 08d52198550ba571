// common.hpp -- shared parts of the command-line drivers.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "epet.h"
#include "workloads.hpp"

namespace epet_drv {

inline void ok(int rc, const char* what) {
  if (rc != EPET_OK)
    throw std::runtime_error(std::string(what) + ": " + epet_last_error());
}

// options after the positional arguments: --workload=NAME (default full)
inline std::string option(int argc, char** argv, int first, const std::string& key,
                          const std::string& dflt) {
  const std::string pre = "--" + key + "=";
  for (int i = first; i < argc; i++) {
    const std::string a(argv[i]);
    if (a.compare(0, pre.size(), pre) == 0) return a.substr(pre.size());
  }
  return dflt;
}

template <typename T>
T* to_device(epet_ctx* ctx, const std::vector<T>& h) {
  void* d = nullptr;
  ok(epet_malloc(ctx, &d, sizeof(T) * (h.empty() ? 1 : h.size())), "epet_malloc");
  if (!h.empty()) ok(epet_memcpy_h2d(ctx, d, h.data(), sizeof(T) * h.size()), "epet_memcpy_h2d");
  return static_cast<T*>(d);
}

template <typename T>
std::vector<T> to_host(epet_ctx* ctx, const T* d, size_t n) {
  std::vector<T> h(n);
  if (n) ok(epet_memcpy_d2h(ctx, h.data(), d, sizeof(T) * n), "epet_memcpy_d2h");
  return h;
}

// FP64 sensitivity over all channels of the scanner (in-tree, replaces the
// sensitivity file): column sums of the nonzero system matrix of every channel
inline std::vector<float> sensitivity_in_tree(epet_ctx* ctx, long long nch, int V) {
  epet_matrix* Afull = nullptr;
  ok(epet_sm_build(ctx, nullptr, (int)nch, &Afull), "epet_sm_build (all channels)");
  double* sd = nullptr;
  ok(epet_malloc(ctx, (void**)&sd, sizeof(double) * V), "epet_malloc");
  ok(epet_matrix_colsum_f64(ctx, Afull, sd), "epet_matrix_colsum_f64");
  std::vector<double> s64 = to_host(ctx, sd, V);
  epet_free(ctx, sd);
  epet_matrix_destroy(Afull);
  return std::vector<float>(s64.begin(), s64.end());
}

}  // namespace epet_drv
