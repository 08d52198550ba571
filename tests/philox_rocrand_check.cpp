// Host-side Philox4x32-10 from the installed open-source rocRAND header, used
// by tests/test_oracle.py as an independent implementation to compare the
// oracle's Philox against.  Reads "c0 c1 c2 c3 k0 k1" lines (hex) from stdin
// and prints the four output words (hex).
#include <rocrand/rocrand_philox4x32_10.h>

#include <cstdio>

struct Philox10 : rocrand_device::philox4x32_10_engine {
  uint4 run(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
  unsigned c0, c1, c2, c3, k0, k1;
  Philox10 p;
  while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
    uint4 c = make_uint4(c0, c1, c2, c3);
    uint2 k = make_uint2(k0, k1);
    uint4 o = p.run(c, k);
    std::printf("%08x %08x %08x %08x\n", o.x, o.y, o.z, o.w);
  }
  return 0;
}
