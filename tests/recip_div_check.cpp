// Exhaustive check of phi_from_count (gpupolyaurnlda_amd/csrc/lda_core.h) against IEEE division:
// every 16-bit significand of the count x (x in [2^15, 2^16)) against every 24-bit significand of
// the divisor d (d in [2^23, 2^24)), plus x = 0 and the S_k = 0 entry (y = 0, d = 1).  The
// identity is invariant under scaling x or d by powers of two in the sampler's range, so this
// covers every count below 2^16 and every topic divisor.  Built with the host compiler
// (-ffp-contract=off: the product and the two fmas stay the rounded operations they are on the
// GPU).  Prints "recip div ok" and exits 0 when no pair differs.
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "lda_core.h"

using namespace lda;

static uint32_t bits(float f) {
  uint32_t b;
  std::memcpy(&b, &f, 4);
  return b;
}

int main(int argc, char** argv) {
  uint32_t d_lo = 1u << 23, d_hi = 1u << 24, step = 1;
  if (argc > 1) step = (uint32_t)std::atoi(argv[1]);  // a stride through the divisors (1: all)
  long long bad = 0, pairs = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : bad, pairs)
  for (uint32_t di = d_lo; di < d_hi; di += step) {
    const float d = (float)di;
    const float y = 1.0f / d;
    for (uint32_t x = 1u << 15; x < (1u << 16); ++x) {
      const float want = (float)x / d;
      bad += bits(phi_from_count(x, y, d)) != bits(want);
    }
    bad += bits(phi_from_count(0, y, d)) != 0u;
    pairs += (1 << 15) + 1;
  }
  bad += bits(phi_from_count(0, 0.0f, 1.0f)) != 0u;
  std::printf("pairs %lld mismatches %lld\n", pairs, bad);
  if (bad) return 1;
  std::printf("recip div ok\n");
  return 0;
}
