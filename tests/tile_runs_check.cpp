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

// The word groups of a slice as the slice build forms them: contiguous word ranges of about
// n / n_groups tokens (a word with more is a group of its own).
static void groups_of(const std::vector<uint64_t>& wstart, uint32_t V, uint64_t n, uint32_t n_groups,
                      std::vector<uint32_t>& gid, std::vector<uint64_t>& gstart) {
  gid.assign(V, 0u);
  gstart.assign(1, 0);
  const uint64_t budget = std::max<uint64_t>(1, (n + n_groups - 1) / n_groups);
  uint64_t acc = 0;
  for (uint32_t v = 0; v < V; ++v) {
    const uint64_t cnt = wstart[v + 1] - wstart[v];
    if (acc > 0 && acc + cnt > budget) {
      gstart.push_back(gstart.back() + acc);
      acc = 0;
    }
    gid[v] = (uint32_t)(gstart.size() - 1);
    acc += cnt;
  }
}

// Both passes built by the parallel per-tile construction, against a sequential walk of the
// outputs: every rank, and every rank's place through the run tables.
static int check_build(uint64_t n, uint32_t V, uint32_t n_groups, unsigned threads, uint64_t seed) {
  std::mt19937_64 rng(seed);
  std::vector<uint32_t> w(n);
  for (auto& x : w) x = (uint32_t)std::min<double>(V - 1, std::floor(V * std::pow((rng() % 1000000 + 1) / 1e6, 3.0)));
  std::vector<uint64_t> wstart(V + 1, 0);
  for (uint32_t x : w) wstart[x + 1]++;
  for (uint32_t v = 0; v < V; ++v) wstart[v + 1] += wstart[v];
  std::vector<uint32_t> gid;
  std::vector<uint64_t> gstart;
  groups_of(wstart, V, n, n_groups, gid, gstart);
  const uint64_t tiles = (n + kTileEntries - 1) / kTileEntries;
  std::vector<uint16_t> rank1, rank2;
  lda::HostRuns r1(tiles), r2(tiles);
  lda::build_reorder_runs(w.data(), n, V, gid.data(), gstart, wstart, rank1, r1, rank2, r2, threads);
  // sequential reference: z1 places p1, word-sorted permutation, ranks by walking the outputs
  std::vector<uint32_t> p1(n), inv1(n), perm(n), want1(n), want2(n), fill(tiles);
  std::vector<uint32_t> place1(tiles * (uint64_t)kTileEntries), place2(tiles * (uint64_t)kTileEntries);
  {
    std::vector<uint64_t> cur(gstart);
    for (uint64_t i = 0; i < n; ++i) {
      p1[i] = (uint32_t)cur[gid[w[i]]]++;
      inv1[p1[i]] = (uint32_t)i;
    }
    std::vector<uint64_t> cw(wstart.begin(), wstart.end() - 1);
    for (uint64_t i = 0; i < n; ++i) perm[cw[w[i]]++] = (uint32_t)i;
  }
  std::fill(fill.begin(), fill.end(), 0u);
  for (uint64_t q = 0; q < n; ++q) {
    const uint32_t i = inv1[q];
    const uint64_t t = i / kTileEntries;
    const uint32_t r = fill[t]++;
    want1[i] = r;
    place1[t * kTileEntries + r] = (uint32_t)q;
  }
  std::fill(fill.begin(), fill.end(), 0u);
  for (uint64_t p = 0; p < n; ++p) {
    const uint32_t q = p1[perm[p]];
    const uint64_t t = q / kTileEntries;
    const uint32_t r = fill[t]++;
    want2[q] = r;
    place2[t * kTileEntries + r] = (uint32_t)p;
  }
  for (uint64_t i = 0; i < n; ++i) {
    if (rank1[i] != want1[i] || rank2[i] != want2[i]) {
      std::fprintf(stderr, "build n=%llu: entry %llu ranks %u/%u, want %u/%u\n", (unsigned long long)n,
                   (unsigned long long)i, rank1[i], rank2[i], want1[i], want2[i]);
      return 1;
    }
  }
  std::vector<uint32_t> pref(kTileWords), off;
  std::vector<lda::RunPair> runs;
  for (int pass = 1; pass <= 2; ++pass) {
    const lda::HostRuns& h = pass == 1 ? r1 : r2;
    const std::vector<uint32_t>& place = pass == 1 ? place1 : place2;
    h.flatten(off, runs);
    for (uint64_t t = 0; t < tiles; ++t) {
      const uint32_t* bits = h.bits.data() + t * kTileWords;
      lda::run_word_prefix(bits, pref.data());
      const uint32_t m = (uint32_t)std::min<uint64_t>(kTileEntries, n - t * kTileEntries);
      for (uint32_t j = 0; j < m; ++j) {
        if (lda::run_place(bits, pref.data(), runs.data() + off[t], j) != place[t * kTileEntries + j]) {
          std::fprintf(stderr, "build n=%llu pass %d: tile %llu rank %u place wrong\n", (unsigned long long)n, pass,
                       (unsigned long long)t, j);
          return 1;
        }
      }
    }
  }
  std::printf("build n=%llu V=%u groups=%zu threads=%u: ranks and places match\n", (unsigned long long)n, V,
              gstart.size(), threads);
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
  for (unsigned th : {1u, 5u})
    for (uint64_t n : {1ull, 1000ull, 32768ull, 100001ull, 400000ull})
      for (uint32_t V : {7u, 300u, 20000u})
        if (check_build(n, V, 256, th, n * 31 + V + th)) return 1;
  if (check_build(300000, 50, 4, 3, 9)) return 1;  // a few large groups, words spanning tiles
  std::printf("tile runs ok\n");
  return 0;
}
