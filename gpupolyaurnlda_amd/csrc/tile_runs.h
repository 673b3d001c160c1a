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

#include <algorithm>
#include <cstdint>
#include <thread>
#include <utility>
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

// fn(lo, hi) over [0, n) in `threads` contiguous parts, each on its own thread.
template <class F>
inline void parallel_ranges(uint64_t n, unsigned threads, F fn) {
  threads = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(threads, n));
  if (threads <= 1) {
    fn((uint64_t)0, n);
    return;
  }
  std::vector<std::thread> pool;
  for (unsigned k = 0; k < threads; ++k) pool.emplace_back(fn, n * k / threads, n * (k + 1) / threads);
  for (auto& th : pool) th.join();
}

// Ranks and output runs of the two reorder passes of a corpus slice of n entries (words w):
//   pass 1: corpus order -> z1 (word groups ascending, corpus order inside a group);
//   pass 2: z1 order -> zs (words ascending, corpus order inside a word).
// gid[v] is the group of word v (non-decreasing in v), gstart[g] the first z1 place of group g,
// wstart[v] the first zs place of word v (V + 1 entries).  rank1[i] is the rank of corpus entry i in
// its tile, rank2[q] that of z1 entry q.  A tile's entries of one group (pass 1) or one word (pass 2)
// are consecutive in the output and keep their order, so each is one run: a tile's ranks are its
// per-group (per-word) counts laid end to end, and the place of a run follows from the counts of
// the tiles before it.  Tiles are independent, so they are built in parallel; the places need one
// ordered scan over per-tile counts.
inline void build_reorder_runs(const uint32_t* w, uint64_t n, uint32_t V, const uint32_t* gid,
                               const std::vector<uint64_t>& gstart, const std::vector<uint64_t>& wstart,
                               std::vector<uint16_t>& rank1, HostRuns& r1, std::vector<uint16_t>& rank2,
                               HostRuns& r2, unsigned threads) {
  const uint64_t tiles = (n + kTileEntries - 1) / kTileEntries;
  const uint32_t G = (uint32_t)gstart.size();
  rank1.assign(n, 0);
  rank2.assign(n, 0);
  auto tile_lo = [&](uint64_t t) { return t * kTileEntries; };
  auto tile_hi = [&](uint64_t t) { return std::min<uint64_t>(n, (t + 1) * kTileEntries); };
  // pass 1: per-tile group counts, their places, then ranks, runs and the words in z1 order
  std::vector<uint32_t> h1(tiles * (uint64_t)G, 0u);
  parallel_ranges(tiles, threads, [&](uint64_t a, uint64_t b) {
    for (uint64_t t = a; t < b; ++t) {
      uint32_t* h = h1.data() + t * G;
      for (uint64_t i = tile_lo(t); i < tile_hi(t); ++i) ++h[gid[w[i]]];
    }
  });
  std::vector<uint64_t> place1(tiles * (uint64_t)G);
  {
    std::vector<uint64_t> cur(gstart);
    for (uint64_t t = 0; t < tiles; ++t)
      for (uint32_t g = 0; g < G; ++g) {
        place1[t * G + g] = cur[g];
        cur[g] += h1[t * G + g];
      }
  }
  std::vector<uint32_t> z1w(n);
  parallel_ranges(tiles, threads, [&](uint64_t a, uint64_t b) {
    std::vector<uint32_t> rp(G), cnt(G);
    for (uint64_t t = a; t < b; ++t) {
      const uint32_t* h = h1.data() + t * G;
      uint32_t acc = 0;
      for (uint32_t g = 0; g < G; ++g) {
        rp[g] = acc;
        cnt[g] = 0;
        if (h[g]) {
          r1.bits[t * kTileWords + acc / 32] |= 1u << (acc % 32);
          r1.per[t].push_back(RunPair{acc, (uint32_t)place1[t * G + g]});
        }
        acc += h[g];
      }
      for (uint64_t i = tile_lo(t); i < tile_hi(t); ++i) {
        const uint32_t g = gid[w[i]];
        const uint32_t c = cnt[g]++;
        rank1[i] = (uint16_t)(rp[g] + c);
        z1w[place1[t * G + g] + c] = w[i];
      }
    }
  });
  std::vector<uint32_t>().swap(h1);
  std::vector<uint64_t>().swap(place1);
  // pass 2: per z1 tile its words (ascending) with counts, their places, then ranks and runs
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> tw(tiles);
  parallel_ranges(tiles, threads, [&](uint64_t a, uint64_t b) {
    std::vector<uint32_t> cnt(V, 0u);
    std::vector<uint32_t> seen;
    for (uint64_t t = a; t < b; ++t) {
      seen.clear();
      for (uint64_t q = tile_lo(t); q < tile_hi(t); ++q) {
        const uint32_t v = z1w[q];
        if (cnt[v]++ == 0) seen.push_back(v);
      }
      std::sort(seen.begin(), seen.end());
      tw[t].reserve(seen.size());
      for (uint32_t v : seen) {
        tw[t].emplace_back(v, cnt[v]);
        cnt[v] = 0;
      }
    }
  });
  std::vector<std::vector<uint32_t>> place2(tiles);
  {
    std::vector<uint64_t> cur(wstart.begin(), wstart.end() - 1);
    for (uint64_t t = 0; t < tiles; ++t) {
      place2[t].resize(tw[t].size());
      for (size_t k = 0; k < tw[t].size(); ++k) {
        place2[t][k] = (uint32_t)cur[tw[t][k].first];
        cur[tw[t][k].first] += tw[t][k].second;
      }
    }
  }
  parallel_ranges(tiles, threads, [&](uint64_t a, uint64_t b) {
    std::vector<uint32_t> rp(V, 0u), cnt(V, 0u);
    for (uint64_t t = a; t < b; ++t) {
      uint32_t acc = 0;
      for (size_t k = 0; k < tw[t].size(); ++k) {
        const uint32_t v = tw[t][k].first;
        rp[v] = acc;
        r2.bits[t * kTileWords + acc / 32] |= 1u << (acc % 32);
        r2.per[t].push_back(RunPair{acc, place2[t][k]});
        acc += tw[t][k].second;
      }
      for (uint64_t q = tile_lo(t); q < tile_hi(t); ++q) {
        const uint32_t v = z1w[q];
        rank2[q] = (uint16_t)(rp[v] + cnt[v]++);
      }
      for (const auto& pr : tw[t]) cnt[pr.first] = 0;
    }
  });
}

}  // namespace lda
