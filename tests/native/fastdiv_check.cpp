// Brute-force check of the index-decomposition division (FastDiv in
// mpf_internal.h): fdiv(n, d) == n / d for every divisor used by grids up to
// 1024 and numerators covering 0..2^24 densely plus random and extreme values.
#include <cstdio>
#include <cstdlib>
#include "../../2017multiphase_amd/csrc/mpf_internal.h"

int main() {
  unsigned long long bad = 0, checked = 0;
  unsigned seed = 12345;
  for (uint32_t d = 1; d <= 1025; ++d) {
    mpf::FastDiv f = mpf::make_fastdiv(d);
    for (uint32_t n = 0; n < (1u << 16); ++n) { bad += mpf::fdiv(n, f) != n / d; ++checked; }
    for (int r = 0; r < 20000; ++r) {
      seed = seed * 1664525u + 1013904223u;
      uint32_t n = seed;
      bad += mpf::fdiv(n, f) != n / d; ++checked;
    }
    const uint32_t ext[] = {0xffffffffu, 0xfffffffeu, 0x80000000u, 0x7fffffffu, d * 1000003u, d - 1, d, d + 1};
    for (uint32_t n : ext) { bad += mpf::fdiv(n, f) != n / d; ++checked; }
  }
  for (uint32_t n = 0; n < (1u << 24); ++n) {
    static const uint32_t ds[] = {16, 24, 64, 96, 128, 256, 257, 384, 385};
    for (uint32_t d : ds) { mpf::FastDiv f = mpf::make_fastdiv(d); bad += mpf::fdiv(n, f) != n / d; ++checked; }
  }
  printf("checked %llu bad %llu\n", checked, bad);
  return bad ? 1 : 0;
}
