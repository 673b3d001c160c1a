// tile_runs.h -- output runs of the n_wk reduction's tile reorders (host side, plain C++: the
// corpus slice build in engine.hip uses it, and a CPU test drives it directly).
//
// A reorder pass moves the entries of each tile of kTileEntries consecutive input entries to their
// places in the pass's output.  Walking the output in order, a tile's entries come up with ranks
// 0, 1, 2, ...; a rank starts a run unless its place follows the previous rank's place.  A tile is
// described by a bitmap of the ranks that start a run (kTileEntries / 32 words) and, per run, its
// first rank and the place of that rank.  The place of rank j is then the place of the run holding
// j plus j's offset in it; the run's index is the number of run starts at ranks <= j, less one
// (rank 0 always starts one), which the device finds from per-word prefix counts of the bitmap.
#pragma once

#include <cstdint>
#include <vector>

namespace lda {

constexpr uint32_t kTileEntries = 32768;
constexpr uint32_t kTileWords = kTileEntries / 32;

struct RunPair {  // same layout as the device's uint2 {x = first rank, y = place}
  uint32_t first;
  uint32_t place;
};

struct HostRuns {
  std::vector<uint32_t> bits;                 // kTileWords words per tile
  std::vector<std::vector<RunPair>> per;      // per tile: its runs in rank order
  std::vector<uint32_t> next;                 // per tile: the place that would continue its last run
  explicit HostRuns(uint64_t n_tiles) : bits(n_tiles * kTileWords, 0u), per(n_tiles), next(n_tiles, 0u) {}
  // Rank r of tile t goes to `place` (ranks of a tile come in order, places increasing).
  void add(uint64_t t, uint32_t r, uint32_t place) {
    if (r == 0 || place != next[t]) {
      bits[t * kTileWords + r / 32] |= 1u << (r % 32);
      per[t].push_back(RunPair{r, place});
    }
    next[t] = place + 1;
  }
  // Runs of all tiles in one array; off[t] .. off[t + 1] are tile t's.
  void flatten(std::vector<uint32_t>& off, std::vector<RunPair>& runs) const {
    off.assign(per.size() + 1, 0u);
    for (size_t t = 0; t < per.size(); ++t) off[t + 1] = off[t] + (uint32_t)per[t].size();
    runs.resize(off.back());
    for (size_t t = 0; t < per.size(); ++t)
      for (size_t k = 0; k < per[t].size(); ++k) runs[off[t] + k] = per[t][k];
  }
};

// Exclusive prefix counts of run starts per bitmap word of one tile.
inline void run_word_prefix(const uint32_t* bits, uint32_t* pref) {
  uint32_t acc = 0;
  for (uint32_t w = 0; w < kTileWords; ++w) {
    pref[w] = acc;
    acc += (uint32_t)__builtin_popcount(bits[w]);
  }
}

// Place of rank j of a tile (bits, pref: the tile's bitmap and prefix counts; runs: its runs), the
// formula of the reorder kernel.
inline uint32_t run_place(const uint32_t* bits, const uint32_t* pref, const RunPair* runs, uint32_t j) {
  const uint32_t w = j >> 5;
  const uint32_t below = bits[w] & (0xFFFFFFFFu >> (31u - (j & 31u)));
  const RunPair r = runs[pref[w] + (uint32_t)__builtin_popcount(below) - 1u];
  return r.place + (j - r.first);
}

}  // namespace lda
