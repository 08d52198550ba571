// backprojection -- back projection of a measurement with the reference's
// command line (tests/backprojection.cu:39-50):
//
//   backprojection <measurement.h5> <output.h5> <nrays> [--workload=NAME] [--seed=S]
//
// x = A^T y over the nonzero channels of dataset "messung", normalised to
// sum 1 (tests/backprojection.cu:115-161), written as dataset "density".
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "common.hpp"
#include "h5io.hpp"

using namespace epet_drv;

int main(int argc, char** argv) {
  if (argc < 4) {
    std::cerr << "Error: Wrong number of arguments. Expected: 3:\n"
                 "  filename of measurement\n  filename of output\n  number of rays\n"
                 "  [--workload=full|small|projector|large] [--seed=S]\n";
    return EXIT_FAILURE;
  }
  const std::string fn(argv[1]), on(argv[2]);
  const int nrays = atoi(argv[3]);
  const std::string wl = option(argc, argv, 4, "workload", "full");
  const unsigned long long seed = std::stoull(option(argc, argv, 4, "seed", "0"));
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
    epet_ctx* ctx = nullptr;
    ok(epet_ctx_create(&ctx, 0, nullptr), "epet_ctx_create");
    ok(epet_set_geometry(ctx, &setup, &grid), "epet_set_geometry");
    ok(epet_set_nrays(ctx, nrays), "epet_set_nrays");
    ok(epet_set_seed(ctx, seed), "epet_set_seed");
    int* cnl_d = to_device(ctx, cnl);
    float* y_d = to_device(ctx, y);
    float* x_d = to_device(ctx, std::vector<float>(V, 0.0f));
    epet_matrix* A = nullptr;
    ok(epet_sm_build(ctx, cnl_d, (int)cnl.size(), &A), "epet_sm_build");
    ok(epet_matrix_spmv(ctx, A, EPET_OP_T, 1.0f, y_d, 0.0f, x_d), "epet_matrix_spmv");
    std::vector<float> x = to_host(ctx, x_d, V);
    double norm = 0.0;
    for (float v : x) norm += v;
    std::cout << "Norm: " << norm << std::endl;
    if (norm != 0.0)
      for (float& v : x) v = (float)(v / norm);
    epet_h5::write_density(on, x.data(), grid);
    epet_matrix_destroy(A);
    epet_free(ctx, cnl_d);
    epet_free(ctx, y_d);
    epet_free(ctx, x_d);
    epet_ctx_destroy(ctx);
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
