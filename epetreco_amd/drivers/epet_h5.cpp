// epet_h5 -- conversions between the reference's HDF5 files and raw
// little-endian float32 files (for scripts without an HDF5 binding):
//
//   epet_h5 mkmeas  <out.h5> <in.f32>                write dataset "messung"
//   epet_h5 mkdens  <out.h5> <in.f32> <workload>     write dataset "density" with attributes
//   epet_h5 dump    <in.h5> <dataset> <out.f32>      read a float dataset
//   epet_h5 attrs   <in.h5>                          print the density attributes
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "h5io.hpp"
#include "workloads.hpp"

static std::vector<float> read_raw(const std::string& fn) {
  std::ifstream f(fn, std::ios::binary | std::ios::ate);
  if (!f) throw std::runtime_error("cannot open " + fn);
  const std::streamsize n = f.tellg();
  f.seekg(0);
  std::vector<float> v(n / 4);
  f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)v.size() * 4);
  return v;
}

static void write_raw(const std::string& fn, const std::vector<float>& v) {
  std::ofstream f(fn, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)v.size() * 4);
}

int main(int argc, char** argv) {
  try {
    const std::string cmd = argc > 1 ? argv[1] : "";
    if (cmd == "mkmeas" && argc == 4) {
      epet_h5::write_measurement(argv[2], read_raw(argv[3]));
    } else if (cmd == "mkdens" && argc == 5) {
      epet_setup s;
      epet_grid g;
      epet_drv::workload(argv[4], &s, &g);
      const std::vector<float> v = read_raw(argv[3]);
      if ((long long)v.size() != (long long)g.n[0] * g.n[1] * g.n[2])
        throw std::runtime_error("density size does not match the grid");
      epet_h5::write_density(argv[2], v.data(), g);
    } else if (cmd == "dump" && argc == 5) {
      write_raw(argv[4], epet_h5::read_floats(argv[2], argv[3]));
    } else if (cmd == "attrs" && argc == 3) {
      const char* fl[6] = {"xmin", "ymin", "zmin", "xmax", "ymax", "zmax"};
      const char* il[3] = {"xnbin", "ynbin", "znbin"};
      for (const char* a : fl) std::printf("%s %.9g\n", a, epet_h5::density_attribute<float>(argv[2], a));
      for (const char* a : il) std::printf("%s %d\n", a, epet_h5::density_attribute<int>(argv[2], a));
    } else {
      std::cerr << "usage: epet_h5 mkmeas|mkdens|dump|attrs ...\n";
      return 2;
    }
  } catch (const std::exception& e) {
    std::cerr << "Error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
