// Drives the run tables of the n_wk reduction's tile reorders (gpupolyaurnlda_amd/csrc/tile_runs.h)
// on the CPU: for output orders of several shapes, every rank's place computed the way the reorder
// kernel computes it (bitmap word prefix + popcount + run table) must be the place the walk gave it.
// Exits nonzero on the first failed check.
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <cmath>
#include <random>
#include <vector>

#include "tile_runs.h"

using lda::kTileEntries;
using lda::kTileWords;

static int check_order(const std::vector<uint32_t>& inv, const char* what) {
  // inv[q] = input entry that goes to output place q
  const uint64_t n = inv.size();
  const uint64_t tiles = (n + kTileEntries - 1) / kTileEntries;
  lda::HostRuns h(tiles);
  std::vector<uint32_t> fill(tiles, 0u), place_of(n);  // place_of[t * kTileEntries + r]
  for (uint64_t q = 0; q < n; ++q) {
    const uint64_t t = inv[q] / kTileEntries;
    const uint32_t r = fill[t]++;
    h.add(t, r, (uint32_t)q);
    place_of[t * kTileEntries + r] = (uint32_t)q;
  }
  std::vector<uint32_t> off;
  std::vector<lda::RunPair> runs;
  h.flatten(off, runs);
  std::vector<uint32_t> pref(kTileWords);
  for (uint64_t t = 0; t < tiles; ++t) {
    const uint32_t m = fill[t];
    if (m != std::min<uint64_t>(kTileEntries, n - t * kTileEntries)) {
      std::fprintf(stderr, "%s: tile %llu holds %u ranks\n", what, (unsigned long long)t, m);
      return 1;
    }
    const uint32_t* bits = h.bits.data() + t * kTileWords;
    lda::run_word_prefix(bits, pref.data());
    for (uint32_t j = 0; j < m; ++j) {
      const uint32_t got = lda::run_place(bits, pref.data(), runs.data() + off[t], j);
      if (got != place_of[t * kTileEntries + j]) {
        std::fprintf(stderr, "%s: tile %llu rank %u: place %u, want %u\n", what, (unsigned long long)t, j, got,
                     place_of[t * kTileEntries + j]);
        return 1;
      }
    }
  }
  std::printf("%s: %llu entries, %llu tiles, %zu runs\n", what, (unsigned long long)n, (unsigned long long)tiles,
              runs.size());
  return 0;
}

int main() {
  std::mt19937_64 rng(7);
  for (uint64_t n : {1ull, 100ull, 32768ull, 32769ull, 200000ull}) {
    std::vector<uint32_t> inv(n);
    std::iota(inv.begin(), inv.end(), 0u);
    if (check_order(inv, "identity")) return 1;  // one run per tile
    std::shuffle(inv.begin(), inv.end(), rng);
    if (check_order(inv, "random")) return 1;  // almost one run per entry
    // word order of a Zipf-like corpus: entries of each word in input order, words ascending
    std::vector<uint32_t> word(n);
    for (uint64_t i = 0; i < n; ++i) word[i] = (uint32_t)std::min<double>(n, 1.0 / std::pow((rng() % 1000000 + 1) / 1e6, 0.9));
    std::iota(inv.begin(), inv.end(), 0u);
    std::stable_sort(inv.begin(), inv.end(), [&](uint32_t a, uint32_t b) { return word[a] < word[b]; });
    if (check_order(inv, "word order")) return 1;
  }
  std::printf("tile runs ok\n");
  return 0;
}
