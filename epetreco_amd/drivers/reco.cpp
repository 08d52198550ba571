// reco -- MLEM reconstruction driver with the reference's command line
// (tests/reco.cu:44-60):
//
//   reco <measurement.h5> <output> <nrays> <sensitivity.h5> <iterations> <guess.h5>
//        [--workload=full|small|projector|large] [--seed=S]
//
// The measurement list is the nonzero channels of dataset "messung"; the
// system matrix of the listed channels is generated once (Philox rays, seed S,
// default 0) instead of every iteration.  The sensitivity file's "density"
// dataset must have one value per voxel; the name "in-tree" computes it (FP64
// column sums over all channels).  A guess file that cannot be opened means
// the homogeneous start 1/V (tests/reco.cu:113-140).  After every iteration the
// image is written to "<it>_<output>" (tests/reco.cu:283-290).
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "common.hpp"
#include "h5io.hpp"

using namespace epet_drv;

int main(int argc, char** argv) {
  if (argc < 7) {
    std::cerr << "Error: Wrong number of arguments. Expected: 6:\n"
                 "  filename of measurement\n  filename of output\n  number of rays\n"
                 "  filename of sensitivity (or in-tree)\n  number reco iterations\n"
                 "  filename of density guess\n"
                 "  [--workload=full|small|projector|large] [--seed=S]\n";
    return EXIT_FAILURE;
  }
  const std::string fn(argv[1]), on(argv[2]), sfn(argv[4]), xfn(argv[6]);
  const int nrays = atoi(argv[3]), nIt = atoi(argv[5]);
  const std::string wl = option(argc, argv, 7, "workload", "full");
  const unsigned long long seed = std::stoull(option(argc, argv, 7, "seed", "0"));
  try {
    epet_setup setup;
    epet_grid grid;
    workload(wl, &setup, &grid);
    const long long nch = nchannels(setup);
    const int V = grid.n[0] * grid.n[1] * grid.n[2];
    std::vector<float> raw = epet_h5::read_measurement(fn);
    if ((long long)raw.size() != nch)
      throw std::runtime_error("measurement has " + std::to_string(raw.size()) +
                               " channels, the scanner " + std::to_string(nch));
    std::vector<int> cnl;
    std::vector<float> y;
    epet_h5::measurement_list(raw, cnl, y);
    std::cout << "measured channels: " << cnl.size() << " of " << nch << std::endl;

    epet_ctx* ctx = nullptr;
    ok(epet_ctx_create(&ctx, 0, nullptr), "epet_ctx_create");
    ok(epet_set_geometry(ctx, &setup, &grid), "epet_set_geometry");
    ok(epet_set_nrays(ctx, nrays), "epet_set_nrays");
    ok(epet_set_seed(ctx, seed), "epet_set_seed");

    std::vector<float> s_host;
    if (sfn == "in-tree") {
      s_host = sensitivity_in_tree(ctx, nch, V);
    } else {
      s_host = epet_h5::read_density(sfn);
      if ((int)s_host.size() != V)
        throw std::runtime_error("Something about the sensitivity file (size?) is wrong");
    }
    std::vector<float> x_host;
    bool xgood = false;
    {
      std::ifstream xf(xfn.c_str());
      xgood = xf.is_open();
    }
    if (xgood) {
      std::cout << "Will use density from file " << xfn << std::endl;
      x_host = epet_h5::read_density(xfn);
      if ((int)x_host.size() != V) throw std::runtime_error("density guess has the wrong size");
    } else {
      std::cout << "No valid density input file given. Will use homogenous density." << std::endl;
      x_host.assign(V, 1.0f / V);
    }

    int* cnl_d = to_device(ctx, cnl);
    float* y_d = to_device(ctx, y);
    float* s_d = to_device(ctx, s_host);
    float* x_d = to_device(ctx, x_host);
    epet_matrix* A = nullptr;
    ok(epet_sm_build(ctx, cnl_d, (int)cnl.size(), &A), "epet_sm_build");
    ok(epet_matrix_transpose(ctx, A), "epet_matrix_transpose");
    for (int it = 0; it < nIt; it++) {
      double ll = 0.0;
      int done = 0;
      ok(epet_mlem(ctx, A, 1, y_d, s_d, x_d, 1, 1, 0.0, &ll, &done), "epet_mlem");
      x_host = to_host(ctx, x_d, V);
      double norm = 0.0;
      for (float v : x_host) norm += v;
      std::cout << "iteration " << it << "  log-likelihood " << ll << "  Norm: " << norm << std::endl;
      epet_h5::write_density(std::to_string(it) + "_" + on, x_host.data(), grid);
    }
    epet_matrix_destroy(A);
    epet_free(ctx, cnl_d);
    epet_free(ctx, y_d);
    epet_free(ctx, s_d);
    epet_free(ctx, x_d);
    epet_ctx_destroy(ctx);
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
