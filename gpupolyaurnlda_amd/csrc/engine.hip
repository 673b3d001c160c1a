// engine.hip -- host side of libGPUPolyaUrnLDA: the five-call ABI of the reference
// (src/cu/train.cu:31-177), the device-resident corpus and the stage entry points.
//
// Device state (one process, one device, not re-entrant -- the reference's contract):
//   phi      f32  V x K   word-major; phi~ after the Poisson draw, phi after normalisation
//   alias    {f32 prob, u32 alias}  V x K word-major
//   sigma_a  f32  V
//   n_wk     u32  V x K   word-major counts (reduced from the sweep)
//   topic_sum u64 K       exact sums of phi~ over words
//   pois     {f32, u32}   100 x 200 Poisson alias tables (lambda = n + beta)
//   segments: device copies of corpus slices (w u32, z u16, doc offsets u64, word-sorted
//             permutation u32, reduction tasks), cached by host pointers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>
#include <thread>
#include <atomic>

#include "../../include/gpulda_abi.h"
#include "../../include/gpulda_ext.h"
#include "tile_runs.h"
#include "kernels.h"
#include "serve_book.h"

using namespace lda;

namespace {

std::string g_err;

int fail(const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

#define HCK(expr)                                                                                \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess) return fail("HIP error: %s %s:%d (%s)", hipGetErrorString(e_), __FILE__, \
                                      __LINE__, #expr);                                          \
  } while (0)

#define CK(expr)            \
  do {                      \
    int r_ = (expr);        \
    if (r_ != 0) return r_; \
  } while (0)

// The ABI calls have no return codes: like the reference (error.cuh:13-18) they print and exit.
void abi_die(int rc) {
  if (rc == 0) return;
  printf("GPUPolyaUrnLDA error: %s\n", g_err.c_str());
  fflush(stdout);
  exit(rc < 0 ? 1 : rc);
}

template <class T>
int dalloc(T** p, u64 n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HCK(hipMalloc((void**)p, n * sizeof(T)));
  return 0;
}

template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

u32 grid_for(u64 n, u32 block) {
  u64 g = (n + block - 1) / block;
  if (g > 65535u * 16) g = 65535u * 16;
  return (u32)std::max<u64>(g, 1);
}

int fill_u32(u32* p, u64 n, u32 v, hipStream_t st) {
  if (n == 0) return 0;
  k_fill_u32<<<grid_for(n, 256), 256, 0, st>>>(p, n, v);
  HCK(hipGetLastError());
  return 0;
}

int fill_u64(unsigned long long* p, u64 n, u64 v, hipStream_t st) {
  if (n == 0) return 0;
  k_fill_u64<<<grid_for(n, 256), 256, 0, st>>>(p, n, (unsigned long long)v);
  HCK(hipGetLastError());
  return 0;
}

int max_lds_per_block() {
  int dev = 0, v = 65536;
  if (hipGetDevice(&dev) == hipSuccess) (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  return v;
}

// ---------------------------------------------------------------------------------------------
// Launchers shared by the ABI, the resident corpus and the stage entry points.
// ---------------------------------------------------------------------------------------------
int lds_per_cu();

int launch_alias_rows(f32* rows, u32 n_rows, u32 m, const unsigned long long* divisor, f32* sigma, f32 alpha,
                      AliasSlot* out, hipStream_t st, u8* wide = nullptr, u8* fit_out = nullptr,
                      const u8* keep_f32 = nullptr, const float2* recip = nullptr) {
  if (n_rows == 0) return 0;
  if (m > kMaxTopics) return fail("alias rows of %u entries (limit %u: slot indices are 16-bit in LDS)", m, kMaxTopics);
  // rows of up to 4096 slots: one workgroup of four wavefronts per row
  if (m <= kAliasBlockThreads * 16 && alias_block_lds(m) <= (u64)max_lds_per_block()) {
    const u64 lds = alias_block_lds(m);
    if (m <= kAliasBlockThreads * 4)
      k_alias_block<4><<<n_rows, kAliasBlockThreads, lds, st>>>(rows, n_rows, m, divisor, sigma, alpha, out,
                                                                wide, fit_out, keep_f32, recip);
    else
      k_alias_block<16><<<n_rows, kAliasBlockThreads, lds, st>>>(rows, n_rows, m, divisor, sigma, alpha, out,
                                                                 wide, fit_out, keep_f32, recip);
    HCK(hipGetLastError());
    return 0;
  }
  if (fit_out) return fail("count rows need the block alias kernel (rows of at most 4096 slots)");
  u64 per = alias_lds_per_wave(m);
  u64 lim = (u64)max_lds_per_block();
  if (per > lim) return fail("row length %u needs %llu B of LDS (limit %llu)", m, (unsigned long long)per,
                             (unsigned long long)lim);
  // wavefronts per block: the count (up to 4) that keeps the most rows resident per CU
  u32 wpb = 1;
  u64 best = 0;
  for (u32 c = 1; c <= 4 && c * per <= lim; ++c) {
    const u64 resident = ((u64)lds_per_cu() / (c * per)) * c;
    if (resident >= best) {
      best = resident;
      wpb = c;
    }
  }
  u32 grid = (n_rows + wpb - 1) / wpb;
  if (m <= 64 * 16) k_alias_rows<16><<<grid, 64 * wpb, per * wpb, st>>>(rows, n_rows, m, divisor, sigma, alpha, out, wpb);
  else k_alias_rows<0><<<grid, 64 * wpb, per * wpb, st>>>(rows, n_rows, m, divisor, sigma, alpha, out, wpb);
  HCK(hipGetLastError());
  return 0;
}

constexpr u32 kPhiRows = 64;

int launch_phi_draw(const u32* n, f32* phi, unsigned long long* tsum, u32 V, u32 K, f32 beta, u32 iter,
                    const u32* iter_ptr, u64 seed, const AliasSlot* tables, hipStream_t st, u16* counts = nullptr,
                    u8* wide = nullptr) {
  CK(fill_u64(tsum, K, 0, st));
  dim3 grid((K + 255) / 256, (V + kPhiRows - 1) / kPhiRows);
  k_phi_draw<<<grid, 256, 0, st>>>(n, phi, tsum, V, K, beta, iter, iter_ptr, (u32)seed, (u32)(seed >> 32), tables,
                                   kPhiRows, counts, wide);
  HCK(hipGetLastError());
  return 0;
}

int lds_per_cu() {
  int dev = 0, v = 65536;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
  return v;
}


// Poisson PMF of lambda = n + beta over 0..199, in double, stored as f32 (poisson.cu:9-20).
void poisson_pmf_rows(f32 beta, std::vector<f32>& pmf) {
  pmf.resize((u64)kPoisTables * kPoisValues);
  for (u32 n = 0; n < kPoisTables; ++n) {
    double lam = (double)n + (double)beta;
    for (u32 x = 0; x < kPoisValues; ++x)
      pmf[(u64)n * kPoisValues + x] = (f32)std::exp((double)x * std::log(lam) - lam - std::lgamma((double)x + 1.0));
  }
}

int build_poisson_tables(f32 beta, AliasSlot* tables, hipStream_t st) {
  std::vector<f32> pmf;
  poisson_pmf_rows(beta, pmf);
  f32* d_pmf = nullptr;
  CK(dalloc(&d_pmf, pmf.size()));
  HCK(hipMemcpyAsync(d_pmf, pmf.data(), pmf.size() * sizeof(f32), hipMemcpyHostToDevice, st));
  int rc = launch_alias_rows(d_pmf, kPoisTables, kPoisValues, nullptr, nullptr, 0.0f, tables, st);
  if (rc == 0) {
    hipError_t e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = fail("HIP error: %s building Poisson tables", hipGetErrorString(e));
  }
  dfree(d_pmf);
  return rc;
}

// ---------------------------------------------------------------------------------------------
// Corpus segments.
// ---------------------------------------------------------------------------------------------
int upload_runs(TileRuns* R, const std::vector<u16>& rank, const HostRuns& h) {
  static_assert(sizeof(RunPair) == sizeof(uint2) && kTileEntries == kCountTile, "run table layout");
  const u64 n = rank.size();
  std::vector<u32> off;
  std::vector<RunPair> runs;
  h.flatten(off, runs);
  CK(dalloc(&R->rank, n));
  CK(dalloc(&R->bits, h.bits.size()));
  CK(dalloc(&R->run_off, off.size()));
  CK(dalloc(&R->runs, runs.size()));
  if (n) HCK(hipMemcpy(R->rank, rank.data(), n * sizeof(u16), hipMemcpyHostToDevice));
  if (!h.bits.empty()) HCK(hipMemcpy(R->bits, h.bits.data(), h.bits.size() * sizeof(u32), hipMemcpyHostToDevice));
  HCK(hipMemcpy(R->run_off, off.data(), off.size() * sizeof(u32), hipMemcpyHostToDevice));
  if (!runs.empty()) HCK(hipMemcpy(R->runs, runs.data(), runs.size() * sizeof(RunPair), hipMemcpyHostToDevice));
  return 0;
}

struct Segment {
  const u32* key_w = nullptr;
  const u32* key_d = nullptr;
  u32 n_docs = 0;
  u64 n_tokens = 0;
  u64 fingerprint = 0;  // content hash of w and d (the cache key with the sizes)
  u64 bytes = 0;        // device bytes this segment holds
  u64 last_use = 0;     // cache clock at its last use (least recently used is evicted first)
  u32 max_len = 0;
  u32 n_tasks = 0;
  u32* w = nullptr;
  u16* z = nullptr;
  u64* doc_off = nullptr;
  // n_wk reduction (launch_count), two tile reorders: z -> z1 (word groups) -> zs (word-sorted);
  // per pass the rank of each input entry inside its kCountTile tile, and the tile's output runs
  // (consecutive ranks going to consecutive places): a bitmap of the ranks that start a run
  // (kCountTile / 32 words per tile) and per run its first rank and place in the output (runs
  // of tile t at run_off[t] .. run_off[t + 1])
  TileRuns runs1, runs2;
  u16* z1 = nullptr;
  u16* zs = nullptr;
  CountTask* tasks = nullptr;
  u32* doc_order = nullptr;  // documents grouped by sweep bucket (longest first inside a bucket)
  u32* doc_lpt = nullptr;    // all documents, longest first (queue of the persistent schedules)
  std::vector<u32> lens_desc;  // document lengths in doc_lpt order (host copy)
  struct Bucket {
    int lmax;      // register-resident lines of the kernel (0 = generic)
    u32 first;     // offset into doc_order
    u32 count;
    u32 max_cap;   // largest map of the bucket (LDS slots)
  };
  std::vector<Bucket> buckets;
  // Pipeline parts of the five-call path: contiguous document ranges of about equal tokens, each
  // swept on its own once its z has arrived, so copies of one part overlap the sweep of another.
  struct Part {
    u32 doc0, doc1;
    u64 tok0, tok1;
    u32 max_len;  // longest document of the part
  };
  std::vector<Part> parts;
  u32* part_lpt = nullptr;  // each part's documents longest first, parts one after another
};

constexpr u32 kMaxParts = 8;
u64 g_part_min_tokens = 4ull << 20;  // smallest pipeline part (gpulda_set_part_tokens)

// n_wk += counts of the segment's (word, topic) pairs: z to word order, then the per-word tasks.
int launch_count(const Segment* sg, u32* n, u32 K, hipStream_t st) {
  if (sg->n_tasks == 0) return 0;
  const u64 tiles = (sg->n_tokens + kCountTile - 1) / kCountTile;
  k_tile_reorder<<<(u32)tiles, 1024, 0, st>>>(sg->z, sg->runs1, sg->z1, sg->n_tokens);
  HCK(hipGetLastError());
  k_tile_reorder<<<(u32)tiles, 1024, 0, st>>>(sg->z1, sg->runs2, sg->zs, sg->n_tokens);
  HCK(hipGetLastError());
  // per wavefront K bins, rounded to 4; up to four wavefronts a block as far as the LDS allows
  const u64 hist = (u64)((K + 3u) & ~3u) * sizeof(u32);
  const u32 wpb = (u32)std::min<u64>(4, (u64)max_lds_per_block() / hist);
  if (wpb == 0) {
    k_count_direct<<<(sg->n_tasks + 3) / 4, 256, 0, st>>>(sg->tasks, sg->n_tasks, sg->zs, n, K);
  } else {
    k_count<<<(sg->n_tasks + wpb - 1) / wpb, 64 * wpb, wpb * hist, st>>>(sg->tasks, sg->n_tasks, sg->zs, n, K, wpb);
  }
  HCK(hipGetLastError());
  return 0;
}

// z-sweep of buckets of documents (order[first ..]): one launch per bucket, the longest maps first.
int launch_buckets(SweepParams P, const std::vector<Segment::Bucket>& buckets, const u32* order, hipStream_t st) {
  const u64 lim = (u64)max_lds_per_block();
  for (const Segment::Bucket& bk : buckets) {
    if (bk.count == 0) continue;
    u64 lds = (u64)doc_counts_bytes(P.K) + (u64)bk.max_cap * 2;  // counts + key table
    if (lds > lim) return fail("doc-topic map of %u slots exceeds LDS", bk.max_cap);
    P.doc_order = order + bk.first;
    switch (bk.lmax) {
      case 1: k_sweep_lines<1><<<bk.count, 64, lds, st>>>(P); break;
      case 2: k_sweep_lines<2><<<bk.count, 64, lds, st>>>(P); break;
      case 4: k_sweep_lines<4><<<bk.count, 64, lds, st>>>(P); break;
      case 8: k_sweep_lines<8><<<bk.count, 64, lds, st>>>(P); break;
      // 16, 32 and larger: maps beyond eight lines take the line-loop draw (the buckets still
      // order the launches and size their LDS)
      default: k_sweep_lines<0><<<bk.count, 64, lds, st>>>(P); break;
    }
    HCK(hipGetLastError());
  }
  return 0;
}

// z-sweep of a whole segment.
int launch_sweep(const SweepParams& P, const Segment* sg, hipStream_t st) {
  return launch_buckets(P, sg->buckets, sg->doc_order, st);
}

void free_segment(Segment* s) {
  if (!s) return;
  dfree(s->w);
  dfree(s->z);
  dfree(s->doc_off);
  for (TileRuns* r : {&s->runs1, &s->runs2}) {
    dfree(r->rank);
    dfree(r->bits);
    dfree(r->run_off);
    dfree(r->runs);
  }
  dfree(s->z1);
  dfree(s->zs);
  dfree(s->tasks);
  dfree(s->doc_order);
  dfree(s->doc_lpt);
  dfree(s->part_lpt);
  delete s;
}

u64 mix64(u64 h, u64 v) {
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h * 0xBF58476D1CE4E5B9ull;
}

// Sampled fingerprint of a slice: all document lengths and 4096 sampled word ids (the trusted
// cache mode, for callers whose w and d never change, such as read-only mappings).
u64 slice_fingerprint(const u32* w, const u32* d, u32 n_docs, u64 n_tokens) {
  u64 h = mix64(n_docs, n_tokens);
  for (u32 i = 0; i < n_docs; ++i) h = mix64(h, d[i]);
  if (n_tokens) {
    u64 step = std::max<u64>(1, n_tokens / 4096);
    for (u64 i = 0; i < n_tokens; i += step) h = mix64(h, w[i]);
    h = mix64(h, w[n_tokens - 1]);
  }
  return h;
}

// 64-bit hash of n u32 words, in chunks of kHashChunk words hashed independently (four
// multiply-rotate lanes each, after xxHash64's round) and combined in chunk order, so it does
// not depend on how the chunks are spread over threads.
constexpr u64 kHashChunk = 1ull << 18;
constexpr u64 kHashP1 = 0x9E3779B185EBCA87ull, kHashP2 = 0xC2B2AE3D27D4EB4Full, kHashP3 = 0x165667B19E3779F9ull;
inline u64 rotl64(u64 x, int r) { return (x << r) | (x >> (64 - r)); }
inline u64 hash_round(u64 acc, u64 x) { return rotl64(acc + x * kHashP2, 31) * kHashP1; }

u64 hash_chunk(const u32* p, u64 n, u64 index) {
  u64 a0 = kHashP1 + index, a1 = kHashP2 ^ index, a2 = kHashP3 + (index << 1), a3 = ~index;
  u64 i = 0;
  for (; i + 8 <= n; i += 8) {
    u64 x[4];
    memcpy(x, p + i, 32);
    a0 = hash_round(a0, x[0]);
    a1 = hash_round(a1, x[1]);
    a2 = hash_round(a2, x[2]);
    a3 = hash_round(a3, x[3]);
  }
  u64 h = rotl64(a0, 1) + rotl64(a1, 7) + rotl64(a2, 12) + rotl64(a3, 18);
  for (; i < n; ++i) h = hash_round(h, p[i]);
  h ^= n;
  h ^= h >> 33;
  h *= kHashP2;
  h ^= h >> 29;
  return h;
}

// Builds device copies of a slice: w, doc offsets, the document queues, the ranks and places of
// the n_wk reduction's two tile reorders and its per-word tasks.  z is allocated, not filled.
int build_segment(Segment* s, const u32* w, const u32* d, u32 n_docs, u64 n_tokens, u32 V, u32 K) {
  std::vector<u64> off((u64)n_docs + 1);
  off[0] = 0;
  u32 max_len = 0;
  for (u32 i = 0; i < n_docs; ++i) {
    if (d[i] > kMaxDocLen) return fail("document %u has %u tokens (limit %u)", i, d[i], kMaxDocLen);
    off[i + 1] = off[i] + d[i];
    max_len = std::max(max_len, d[i]);
  }
  if (off[n_docs] != n_tokens)
    return fail("document lengths sum to %llu, expected %llu tokens", (unsigned long long)off[n_docs],
                (unsigned long long)n_tokens);
  std::vector<u64> start((u64)V + 1, 0);
  for (u64 i = 0; i < n_tokens; ++i) {
    if (w[i] >= V) return fail("word id %u at token %llu is >= V = %u", w[i], (unsigned long long)i, V);
    start[(u64)w[i] + 1]++;
  }
  std::vector<CountTask> tasks;
  for (u32 v = 0; v < V; ++v) {
    u64 cnt = start[(u64)v + 1];
    start[(u64)v + 1] = start[v] + cnt;
    for (u64 p = 0; p < cnt; p += kCountTask) {
      CountTask t;
      t.p0 = start[v] + p;
      t.len = (u32)std::min<u64>(kCountTask, cnt - p);
      t.word = v;
      tasks.push_back(t);
    }
  }
  if (n_tokens >= (1ull << 32)) return fail("a slice must hold fewer than 2^32 tokens");
  // Ranks and destinations of the two tile reorders of the n_wk reduction.  Word groups are
  // contiguous word ranges of about n_tokens / kCountGroups tokens (a word with more is a group of
  // its own).  z1 order: groups ascending, corpus order inside a group; word-sorted order (zs):
  // words ascending, corpus order inside a word.  Built tile by tile on host threads
  // (build_reorder_runs, tile_runs.h).
  const u64 n_tiles = (n_tokens + kCountTile - 1) / kCountTile;
  std::vector<u16> crank1, crank2;
  HostRuns hr1(n_tiles), hr2(n_tiles);
  {
    std::vector<u32> gid(V);
    std::vector<u64> gstart(1, 0);
    const u64 budget = std::max<u64>(1, (n_tokens + kCountGroups - 1) / kCountGroups);
    u64 acc = 0;
    for (u32 v = 0; v < V; ++v) {
      const u64 cnt = start[(u64)v + 1] - start[v];
      if (acc > 0 && acc + cnt > budget) {
        gstart.push_back(gstart.back() + acc);
        acc = 0;
      }
      gid[v] = (u32)(gstart.size() - 1);
      acc += cnt;
    }
    const unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    build_reorder_runs(w, n_tokens, V, gid.data(), gstart, start, crank1, hr1, crank2, hr2, threads);
  }
  // sweep buckets: by register-resident line count, largest first; longest documents first
  std::vector<u32> order(n_docs);
  std::vector<int> lmax_of(n_docs);
  for (u32 i = 0; i < n_docs; ++i) {
    order[i] = i;
    lmax_of[i] = sweep_bucket_lmax(map_capacity(d[i], K) / kMapLine);
  }
  auto rank = [](int lm) { return lm == 0 ? 1000 : lm; };
  std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) {
    if (lmax_of[a] != lmax_of[b]) return rank(lmax_of[a]) > rank(lmax_of[b]);
    return d[a] > d[b];
  });
  s->buckets.clear();
  for (u32 i = 0; i < n_docs; ++i) {
    u32 doc = order[i];
    u32 cap = map_capacity(d[doc], K);
    if (s->buckets.empty() || s->buckets.back().lmax != lmax_of[doc]) s->buckets.push_back({lmax_of[doc], i, 0, 0});
    s->buckets.back().count++;
    s->buckets.back().max_cap = std::max(s->buckets.back().max_cap, cap);
  }
  // pipeline parts: about n_tokens / kMaxParts tokens each (at least g_part_min_tokens), whole
  // documents, each part's documents longest first
  s->parts.clear();
  std::vector<u32> plpt(n_docs);
  {
    const u64 part_tokens = std::max<u64>(std::max<u64>(1, g_part_min_tokens), (n_tokens + kMaxParts - 1) / kMaxParts);
    u32 doc = 0;
    u64 tok = 0;
    while (doc < n_docs) {
      Segment::Part pt;
      pt.doc0 = doc;
      pt.tok0 = tok;
      while (doc < n_docs && (doc == pt.doc0 || tok - pt.tok0 < part_tokens)) tok += d[doc++];
      pt.doc1 = doc;
      pt.tok1 = tok;
      for (u32 i = pt.doc0; i < pt.doc1; ++i) plpt[i] = i;
      std::stable_sort(plpt.begin() + pt.doc0, plpt.begin() + pt.doc1, [&](u32 a, u32 b) { return d[a] > d[b]; });
      pt.max_len = d[plpt[pt.doc0]];
      s->parts.push_back(pt);
    }
  }
  CK(dalloc(&s->part_lpt, n_docs));
  if (n_docs) HCK(hipMemcpy(s->part_lpt, plpt.data(), (u64)n_docs * sizeof(u32), hipMemcpyHostToDevice));
  s->n_docs = n_docs;
  s->n_tokens = n_tokens;
  s->max_len = max_len;
  s->n_tasks = (u32)tasks.size();
  CK(dalloc(&s->doc_order, n_docs));
  if (n_docs) HCK(hipMemcpy(s->doc_order, order.data(), (u64)n_docs * sizeof(u32), hipMemcpyHostToDevice));
  std::vector<u32> lpt(n_docs);
  for (u32 i = 0; i < n_docs; ++i) lpt[i] = i;
  std::stable_sort(lpt.begin(), lpt.end(), [&](u32 a, u32 b) { return d[a] > d[b]; });
  s->lens_desc.resize(n_docs);
  for (u32 i = 0; i < n_docs; ++i) s->lens_desc[i] = d[lpt[i]];
  CK(dalloc(&s->doc_lpt, n_docs));
  if (n_docs) HCK(hipMemcpy(s->doc_lpt, lpt.data(), (u64)n_docs * sizeof(u32), hipMemcpyHostToDevice));
  CK(dalloc(&s->w, n_tokens));
  CK(dalloc(&s->z, n_tokens));
  CK(dalloc(&s->doc_off, (u64)n_docs + 1));
  CK(upload_runs(&s->runs1, crank1, hr1));
  CK(upload_runs(&s->runs2, crank2, hr2));
  CK(dalloc(&s->z1, n_tokens));
  CK(dalloc(&s->zs, n_tokens));
  CK(dalloc(&s->tasks, tasks.size()));
  if (n_tokens) {
    HCK(hipMemcpy(s->w, w, n_tokens * sizeof(u32), hipMemcpyHostToDevice));
  }
  HCK(hipMemcpy(s->doc_off, off.data(), off.size() * sizeof(u64), hipMemcpyHostToDevice));
  if (!tasks.empty()) HCK(hipMemcpy(s->tasks, tasks.data(), tasks.size() * sizeof(CountTask), hipMemcpyHostToDevice));
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Working memory of the allocation-free entry point (gpulda_foldin_noalloc): one static device
// array per GPU, held until the process ends.  Host code never uses the array's name as an address
// (on the host it is only a stand-in); its device address on each GPU comes from
// hipGetSymbolAddress, taken once per device with that device current and kept here.
// ---------------------------------------------------------------------------------------------
constexpr u64 kNoAllocScratchBytes = 16ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on its first use there).
int scratch_on_current_device(unsigned char** out) {
  int dev = 0;
  HCK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail("device %d beyond the %d this library tracks", dev, kMaxDevices);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HCK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail("no device address for the working-memory array");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Library state.
// ---------------------------------------------------------------------------------------------
struct BufferState {
  hipStream_t stream = nullptr;    // sweeps and counts of this buffer
  hipStream_t copy_in = nullptr;   // host -> device copies of z
  hipStream_t copy_out = nullptr;  // device -> host copies of z
  hipEvent_t ev_in[kMaxParts] = {};
  hipEvent_t ev_swept[kMaxParts] = {};
  hipEvent_t ev_out[kMaxParts] = {};
  hipEvent_t ev_done = nullptr;
  u16* host_z = nullptr;   // pinned u16 z of the slice in flight (buffer_size)
  u32 host_bad = 0;        // indicators >= K seen while narrowing (reported by sync_buffer)
  u32* heads = nullptr;    // queue heads of the parts' persistent sweeps (kMaxParts)
  Segment* seg = nullptr;  // segment of the last call (uncached mode owns it)
  // work left for sync_buffer: widen the parts' topics into the caller's z as they arrive
  Segment* pending = nullptr;
  u32* pending_z = nullptr;
  Buffer* owner = nullptr;  // the caller's struct (its gpu_w / gpu_d_len name the segment in use)
};

// Threads for host-side conversions of z (the caller's thread takes the first share).
u32 host_threads() {
  const u32 hc = std::thread::hardware_concurrency();
  return std::max<u32>(1, std::min<u32>(16, hc ? hc : 1));
}

// f(begin, end) over [0, n), split across host_threads() threads (n of at least min_split).
template <class F>
void host_parallel(u64 n, F f, u64 min_split = 1ull << 16) {
  const u32 T = host_threads();
  if (T == 1 || n < min_split) {
    f((u64)0, n);
    return;
  }
  const u64 chunk = (n + T - 1) / T;
  std::vector<std::thread> th;
  for (u32 t = 1; t < T; ++t) {
    const u64 a = (u64)t * chunk, b = std::min<u64>(n, a + chunk);
    if (a < b) th.emplace_back(f, a, b);
  }
  f((u64)0, std::min<u64>(n, chunk));
  for (std::thread& x : th) x.join();
}

// u32 topics -> u16 (an indicator >= K becomes topic 0 and is counted).
u32 narrow_topics(const u32* in, u16* out, u64 n, u32 K) {
  std::atomic<u32> bad{0};
  host_parallel(n, [&](u64 a, u64 b) {
    u32 nb = 0;
    for (u64 i = a; i < b; ++i) {
      const u32 v = in[i];
      const bool ok = v < K;
      nb += !ok;
      out[i] = (u16)(ok ? v : 0u);
    }
    bad += nb;
  });
  return bad.load();
}

void widen_topics(const u16* in, u32* out, u64 n) {
  host_parallel(n, [&](u64 a, u64 b) {
    for (u64 i = a; i < b; ++i) out[i] = in[i];
  });
}

// Hash of a slice's whole content: its sizes, every document length and every word id.
u64 slice_content_hash(const u32* w, const u32* d, u32 n_docs, u64 n_tokens) {
  const u64 nc = (n_tokens + kHashChunk - 1) / kHashChunk;
  std::vector<u64> part(nc);
  host_parallel(
      nc,
      [&](u64 a, u64 b) {
        for (u64 c = a; c < b; ++c)
          part[c] = hash_chunk(w + c * kHashChunk, std::min(kHashChunk, n_tokens - c * kHashChunk), c);
      },
      2);
  u64 h = mix64(mix64(n_docs, n_tokens), hash_chunk(d, n_docs, ~0ull));
  for (u64 x : part) h = mix64(h, x);
  return h;
}

enum Schedule : int { kSchedBuckets = 0, kSchedPersistent = 1, kSchedPersistentBlock = 2, kSchedPersistentSplit = 3 };

struct State {
  bool ready = false;
  Args* args = nullptr;
  u32 K = 0, V = 0;
  f32 alpha = 0, beta = 0;
  u64 seed = 0;
  u32 iter = 0;
  u32 doc_counter = 0;
  // device copies of buffer slices: 0 = none (every call uploads, as the reference), 1 = reused
  // when the whole content matches (default), 2 = reused when the host pointers, sizes and a
  // sampled fingerprint match (for callers whose w and d never change)
  int slice_cache = 1;
  u64 cache_budget = 0;       // bytes the cached slices may hold (0: half the free HBM at initialize)
  u64 cache_auto_budget = 0;
  u64 cache_bytes = 0;
  u64 cache_clock = 0;
  u64 cache_builds = 0;       // slices built (uploads of w and d)
  std::vector<BufferState*> buffers;
  hipStream_t stream = nullptr;  // Phi stream / resident stream
  f32* phi = nullptr;
  AliasSlot* alias = nullptr;
  f32* sigma_a = nullptr;
  // u16 rows of the phi~ counts beside the f32 rows, written by the Phi step (count_rows on):
  // the persistent sweep reads them for the words whose counts fit (half the bytes per row)
  bool count_rows = true;
  u16* phi_counts = nullptr;
  u8* counts_fit = nullptr;  // V, padded to 4 bytes
  u8* keep_f32 = nullptr;    // V, padded to 4 bytes: the most frequent words, read as f32 rows
  u8* counts_wide = nullptr; // V, padded to 4 bytes: a count of the row passed 16 bits (per Phi step)
  std::vector<u32> host_C;   // corpus counts per word given to initialize (empty: none given)
  float2* topic_recip = nullptr;
  u32* n = nullptr;
  u32* C = nullptr;
  unsigned long long* topic_sum = nullptr;
  unsigned long long* visited = nullptr;
  AliasSlot* pois = nullptr;
  std::vector<Segment*> cache;
  Segment* resident = nullptr;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  u32* iter_dev = nullptr;   // iteration counter on the device (bumped at each Phi step)
  u32* heads = nullptr;      // work-queue heads of the persistent sweeps
  // resident-corpus sweep schedule (see launch_sweep_schedule)
  int schedule = kSchedBuckets;
  u32 waves_per_cu = 16;
  u32 block_min_len = 0;
  bool use_graph = false;
  hipGraphExec_t graph = nullptr;
  hipStream_t stream2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  int cu_count = 0;
};

State S;
u64 g_seed = 0;

// A captured iteration refers to the resident segment and the schedule: rebuilt on change.
void drop_graph() {
  if (S.graph) (void)hipGraphExecDestroy(S.graph);
  S.graph = nullptr;
}

int require_ready() {
  if (!S.ready) return fail("library not initialised (call initialize first)");
  return 0;
}

// Phi step of the next iteration: bump the device iteration counter, Poisson draw, alias rows,
// and clear n_wk for the sweeps that follow (sample_phi, train.cu:137-156).
int phi_step(hipStream_t st, hipEvent_t mid = nullptr) {
  S.iter += 1;
  S.doc_counter = 0;
  k_bump_u32<<<1, 64, 0, st>>>(S.iter_dev);
  HCK(hipGetLastError());
  if (S.count_rows)
    CK(launch_phi_draw(S.n, S.phi, S.topic_sum, S.V, S.K, S.beta, 0, S.iter_dev, S.seed, S.pois, st, S.phi_counts,
                       S.counts_wide));
  else
    CK(launch_phi_draw(S.n, S.phi, S.topic_sum, S.V, S.K, S.beta, 0, S.iter_dev, S.seed, S.pois, st));
  if (mid) HCK(hipEventRecord(mid, st));
  k_topic_recip<<<(S.K + 1 + 255) / 256, 256, 0, st>>>(S.topic_sum, S.topic_recip, S.K);
  HCK(hipGetLastError());
  if (S.count_rows)
    CK(launch_alias_rows(S.phi, S.V, S.K, S.topic_sum, S.sigma_a, S.alpha, S.alias, st, S.counts_wide,
                         S.counts_fit, S.keep_f32, S.topic_recip));
  else
    CK(launch_alias_rows(S.phi, S.V, S.K, S.topic_sum, S.sigma_a, S.alpha, S.alias, st, nullptr, nullptr, nullptr,
                         S.topic_recip));
  CK(fill_u32(S.n, (u64)S.V * S.K, 0, st));
  return 0;
}

SweepParams sweep_params(Segment* sg, u32 doc_base) {
  SweepParams P;
  P.w = sg->w;
  P.z = sg->z;
  P.doc_off = sg->doc_off;
  P.doc_order = nullptr;
  P.n_docs = sg->n_docs;
  P.doc_base = doc_base;
  P.iter = 0;
  P.iter_ptr = S.iter_dev;
  P.k0 = (u32)S.seed;
  P.k1 = (u32)(S.seed >> 32);
  P.K = S.K;
  P.phi = S.phi;
  P.sigma_a = S.sigma_a;
  P.alias = S.alias;
  P.visited = S.visited;
  if (S.count_rows) {
    P.phi_counts = S.phi_counts;
    P.counts_fit = S.counts_fit;
    P.topic_recip = S.topic_recip;
  }
  return P;
}

// LDS bytes one wavefront needs for any document of the segment (counts + key table).
u64 sweep_lds_bytes(const Segment* sg) { return doc_state_bytes(sg->max_len, S.K); }

// Persistent sweep of documents whose maps stay within the register variants (k_sweep_persistent<8>),
// at up to waves_per_cu wavefronts per CU as LDS (lds bytes per wavefront) allows.  With count rows
// it runs k_sweep_persistent_counts in blocks of up to kCountsBlockWaves wavefronts sharing the
// topic reciprocals, when that keeps at least as many wavefronts per CU.
int launch_persistent_short(const SweepParams& P, u32* head, u32 n, u64 lds, u32 waves_per_cu, hipStream_t st) {
  if (n == 0) return 0;
  const u32 per_cu = std::max<u32>(1, std::min<u32>(waves_per_cu, (u32)(lds_per_cu() / lds)));
  if (P.phi_counts) {
    const u32 wb = std::min<u32>(kCountsBlockWaves, per_cu);
    const u64 blk = topic_recip_lds_bytes(P.K) + (u64)wb * lds;
    const u32 blocks_per_cu = std::min<u32>((u32)(lds_per_cu() / blk), (per_cu + wb - 1) / wb);
    if (blk <= (u64)max_lds_per_block() && blocks_per_cu * wb >= per_cu) {
      const u32 blocks = std::min<u32>((u32)S.cu_count * blocks_per_cu, (n + wb - 1) / wb);
      k_sweep_persistent_counts<<<blocks, 64 * wb, blk, st>>>(P, head, n, (u32)lds);
      HCK(hipGetLastError());
      return 0;
    }
  }
  const u32 waves = std::min<u32>((u32)S.cu_count * per_cu, n);
  k_sweep_persistent<8><<<waves, 64, lds, st>>>(P, head, n);
  HCK(hipGetLastError());
  return 0;
}

// Resident sweep under the configured schedule:
//   buckets              one launch per length bucket (k_sweep_lines<LMAX>), longest maps first;
//   persistent           one launch of cu_count * waves_per_cu wavefronts taking documents from a
//                        longest-first queue (k_sweep_persistent);
//   persistent + block   documents of at least block_min_len tokens go to a block-per-document
//                        persistent launch (k_sweep_block) on a second stream, the rest as above;
//   persistent split     documents longer than block_min_len go to a persistent launch of the
//                        line-loop sweep (k_sweep_persistent_long, few registers: long documents
//                        grow maps past the register variants) sized for the longest map, on a
//                        second stream; the others to k_sweep_persistent<8> whose LDS is sized
//                        for their own longest map, so more of its wavefronts fit on a CU.
// All schedules produce bit-identical z: a document's randomness depends on its global id only.
int launch_sweep_schedule(const SweepParams& P0, Segment* sg, hipStream_t st) {
  if (S.schedule == kSchedBuckets || sg->n_docs == 0) return launch_sweep(P0, sg, st);
  SweepParams P = P0;
  P.doc_order = sg->doc_lpt;
  const u64 lds = sweep_lds_bytes(sg);
  if (lds > (u64)max_lds_per_block()) return fail("doc-topic map exceeds LDS");
  // Long documents and the rest run side by side: the long documents' launch goes on the launch
  // stream, the rest on the second stream behind an event, so that the long documents' wavefronts
  // are placed first (a long document is one serial chain; started behind a persistent launch
  // that holds every slot, it ends that much later).
  u32 n_long = 0;
  if (S.schedule == kSchedPersistentSplit)
    while (n_long < sg->n_docs && sg->lens_desc[n_long] > S.block_min_len) ++n_long;
  else if (S.schedule == kSchedPersistentBlock && S.block_min_len > 0)
    while (n_long < sg->n_docs && sg->lens_desc[n_long] >= S.block_min_len) ++n_long;
  CK(fill_u32(S.heads, 2, 0, st));
  const bool both = n_long > 0 && n_long < sg->n_docs;
  if (both) {
    HCK(hipEventRecord(S.fork, st));
    HCK(hipStreamWaitEvent(S.stream2, S.fork, 0));
  }
  if (n_long > 0) {
    if (S.schedule == kSchedPersistentSplit) {
      const u32 per_cu = std::max<u32>(1, std::min<u32>(S.waves_per_cu, (u32)(lds_per_cu() / lds)));
      const u32 waves = std::min<u32>((u32)S.cu_count * per_cu, n_long);
      k_sweep_persistent_long<<<waves, 64, lds, st>>>(P, S.heads + 1, n_long);
    } else {
      const u32 blocks = std::min<u32>(n_long, (u32)S.cu_count * 2);
      k_sweep_block<<<blocks, 256, lds, st>>>(P, S.heads + 1, n_long);
    }
    HCK(hipGetLastError());
  }
  if (n_long < sg->n_docs) {
    SweepParams Q = P;
    Q.doc_order = sg->doc_lpt + n_long;
    const u64 lds_s = S.schedule == kSchedPersistentSplit ? doc_state_bytes(sg->lens_desc[n_long], S.K) : lds;
    hipStream_t rs = both ? S.stream2 : st;
    CK(launch_persistent_short(Q, S.heads, sg->n_docs - n_long, lds_s, S.waves_per_cu, rs));
    if (both) {
      HCK(hipEventRecord(S.join, S.stream2));
      HCK(hipStreamWaitEvent(st, S.join, 0));
    }
  }
  return 0;
}

// z-sweep of one pipeline part of a five-call slice: one persistent launch over the part's
// documents, longest first (head: the part's queue counter, zero), at kAbiWavesPerCu wavefronts
// per CU or as many as LDS holds -- the schedule the autotuner picks for the resident corpus on
// MI355X.
constexpr u32 kAbiWavesPerCu = 32;
int launch_part_sweep(const SweepParams& P0, const Segment* sg, u32 p, u32* head, hipStream_t st) {
  const Segment::Part& pt = sg->parts[p];
  const u32 n = pt.doc1 - pt.doc0;
  if (n == 0) return 0;
  SweepParams P = P0;
  P.doc_order = sg->part_lpt + pt.doc0;
  const u64 lds = doc_state_bytes(pt.max_len, S.K);
  if (lds > (u64)max_lds_per_block()) return fail("doc-topic map exceeds LDS");
  return launch_persistent_short(P, head, n, lds, kAbiWavesPerCu, st);
}

// A segment some buffer still has in flight (between its sample_z_async and sync_buffer).
bool segment_pinned(const Segment* s) {
  for (const BufferState* b : S.buffers)
    if (b->pending == s) return true;
  return false;
}

u64 segment_bytes(const Segment* s, u64 n_tasks) {
  // w, z, and per count pass rank, destination and output (z1, zs) per token
  return s->n_tokens * (sizeof(u32) + sizeof(u16) + 2 * (sizeof(u16) + sizeof(u32) + sizeof(u16))) +
         ((u64)s->n_docs + 1) * sizeof(u64) +
         (u64)s->n_docs * 3 * sizeof(u32) + n_tasks * sizeof(CountTask);
}

// Frees cached segment i.  No work is queued on it: it is not pinned, and a buffer's work ends at
// its sync_buffer.  Callers' Buffer structs that still name it forget it.
void evict_segment(size_t i) {
  Segment* s = S.cache[i];
  for (BufferState* b : S.buffers)
    if (b->owner && b->owner->gpu_w == (void*)s->w) {
      b->owner->gpu_w = nullptr;
      b->owner->gpu_d_len = nullptr;
    }
  S.cache_bytes -= s->bytes;
  free_segment(s);
  S.cache.erase(S.cache.begin() + (long)i);
}

// Evicts the least recently used segments that no buffer has in flight until `need` more bytes
// fit the budget (or only pinned segments are left).
void make_room(u64 need) {
  const u64 budget = S.cache_budget ? S.cache_budget : S.cache_auto_budget;
  while (!S.cache.empty() && S.cache_bytes + need > budget) {
    long victim = -1;
    for (size_t i = 0; i < S.cache.size(); ++i)
      if (!segment_pinned(S.cache[i]) && (victim < 0 || S.cache[i]->last_use < S.cache[(size_t)victim]->last_use))
        victim = (long)i;
    if (victim < 0) return;
    evict_segment((size_t)victim);
  }
}

int get_segment(BufferState* bs, const Buffer* B, Segment** out) {
  const u32* w = B->w;
  const u32* d = B->d;
  if (S.slice_cache == 0) {
    HCK(hipStreamSynchronize(bs->stream));
    free_segment(bs->seg);
    bs->seg = new Segment();
    ++S.cache_builds;
    CK(build_segment(bs->seg, w, d, B->n_docs, B->n_tokens, S.V, S.K));
    *out = bs->seg;
    return 0;
  }
  const bool trusted = S.slice_cache == 2;
  const u64 fp = trusted ? slice_fingerprint(w, d, B->n_docs, B->n_tokens)
                         : slice_content_hash(w, d, B->n_docs, B->n_tokens);
  const u32* key_w = trusted ? w : nullptr;  // a content-verified entry is not tied to the pointers
  const u32* key_d = trusted ? d : nullptr;
  for (Segment* s : S.cache)
    if (s->key_w == key_w && s->key_d == key_d && s->n_docs == B->n_docs && s->n_tokens == B->n_tokens &&
        s->fingerprint == fp && !segment_pinned(s)) {  // (another buffer in flight on it: a copy of its own)
      s->last_use = ++S.cache_clock;
      *out = s;
      return 0;
    }
  Segment probe;
  probe.n_docs = B->n_docs;
  probe.n_tokens = B->n_tokens;
  make_room(segment_bytes(&probe, (u64)S.V + B->n_tokens / kCountTask + 1));
  Segment* s = new Segment();
  ++S.cache_builds;
  int rc = build_segment(s, w, d, B->n_docs, B->n_tokens, S.V, S.K);
  if (rc != 0 && g_err.find("out of memory") != std::string::npos) {
    // evict every slice no buffer has in flight and retry once
    free_segment(s);
    HCK(hipDeviceSynchronize());
    for (size_t i = S.cache.size(); i-- > 0;)
      if (!segment_pinned(S.cache[i])) evict_segment(i);
    s = new Segment();
    rc = build_segment(s, w, d, B->n_docs, B->n_tokens, S.V, S.K);
  }
  if (rc != 0) {
    free_segment(s);
    return rc;
  }
  s->key_w = key_w;
  s->key_d = key_d;
  s->fingerprint = fp;
  s->bytes = segment_bytes(s, s->n_tasks);
  s->last_use = ++S.cache_clock;
  S.cache_bytes += s->bytes;
  S.cache.push_back(s);
  *out = s;
  return 0;
}

void release_all() {
  if (S.stream) (void)hipDeviceSynchronize();
  for (Segment* s : S.cache) free_segment(s);
  S.cache.clear();
  S.cache_bytes = 0;
  S.buffers.clear();
  free_segment(S.resident);
  S.resident = nullptr;
  dfree(S.phi);
  dfree(S.alias);
  dfree(S.sigma_a);
  dfree(S.phi_counts);
  dfree(S.counts_fit);
  dfree(S.keep_f32);
  dfree(S.counts_wide);
  dfree(S.topic_recip);
  dfree(S.n);
  dfree(S.C);
  dfree(S.topic_sum);
  dfree(S.visited);
  dfree(S.pois);
  dfree(S.iter_dev);
  dfree(S.heads);
  if (S.graph) (void)hipGraphExecDestroy(S.graph);
  S.graph = nullptr;
  if (S.fork) (void)hipEventDestroy(S.fork);
  if (S.join) (void)hipEventDestroy(S.join);
  S.fork = S.join = nullptr;
  if (S.stream2) (void)hipStreamDestroy(S.stream2);
  S.stream2 = nullptr;
  for (auto& e : S.ev)
    if (e) {
      (void)hipEventDestroy(e);
      e = nullptr;
    }
  if (S.stream) (void)hipStreamDestroy(S.stream);
  S.stream = nullptr;
  S.ready = false;
}

// The n most frequent words (by the corpus counts C of initialize) keep f32 rows in the sweep.
// Default: as many as 1 MiB of f32 rows, a quarter of one XCD's L2 -- their rows stay in L2, where
// the count rows' conversion would cost issue slots and save no misses (large workload: 256
// words, 61 % of the tokens; the sweep 58.2 -> 56.3 ms, flat from 100 to 300 words).
int set_keep_f32(u32 n_keep) {
  std::vector<u8> keep(((u64)S.V + 3) / 4 * 4, 0);
  const u32 r = std::min<u32>(n_keep, S.host_C.empty() ? 0u : S.V);
  if (r > 0) {
    std::vector<u32> ids(S.V);
    for (u32 v = 0; v < S.V; ++v) ids[v] = v;
    const std::vector<u32>& C = S.host_C;
    std::nth_element(ids.begin(), ids.begin() + (r - 1), ids.end(),
                     [&](u32 a, u32 b) { return C[a] != C[b] ? C[a] > C[b] : a < b; });
    for (u32 i = 0; i < r; ++i) keep[ids[i]] = 1;
  }
  HCK(hipMemcpy(S.keep_f32, keep.data(), keep.size(), hipMemcpyHostToDevice));
  return 0;
}

int do_initialize(Args* args, Buffer* buffers, u32 n_buffers) {
  if (!args) return fail("initialize: args is NULL");
  gpulda_serve_close();
  if (S.ready) release_all();
  if (args->K == 0 || args->K > 4096) return fail("K = %u outside 1..4096", args->K);
  if (args->V == 0) return fail("V must be positive");
  if (!(args->alpha > 0.0f) || !(args->beta > 0.0f)) return fail("alpha and beta must be positive");
  if (alias_lds_per_wave(args->K) > (u64)max_lds_per_block())
    return fail("K = %u needs more LDS per alias row than a workgroup has", args->K);
  S.args = args;
  S.K = args->K;
  S.V = args->V;
  S.alpha = args->alpha;
  S.beta = args->beta;
  S.seed = g_seed;
  S.iter = 0;
  S.doc_counter = 0;
  HCK(hipStreamCreate(&S.stream));
  HCK(hipStreamCreate(&S.stream2));
  HCK(hipEventCreateWithFlags(&S.fork, hipEventDisableTiming));
  HCK(hipEventCreateWithFlags(&S.join, hipEventDisableTiming));
  for (auto& e : S.ev) HCK(hipEventCreate(&e));
  {
    int dev = 0;
    HCK(hipGetDevice(&dev));
    HCK(hipDeviceGetAttribute(&S.cu_count, hipDeviceAttributeMultiprocessorCount, dev));
  }
  {
    unsigned char* scr = nullptr;
    CK(scratch_on_current_device(&scr));  // set-up of the allocation-free entry point
  }
  CK(dalloc(&S.iter_dev, 1));
  CK(dalloc(&S.heads, 2));
  CK(fill_u32(S.iter_dev, 1, 0, S.stream));
  const u64 VK = (u64)S.V * S.K;
  CK(dalloc(&S.phi, VK));
  CK(dalloc(&S.alias, VK));
  CK(dalloc(&S.sigma_a, S.V));
  CK(dalloc(&S.phi_counts, VK));
  CK(dalloc(&S.counts_fit, ((u64)S.V + 3) / 4 * 4));
  CK(dalloc(&S.keep_f32, ((u64)S.V + 3) / 4 * 4));
  CK(dalloc(&S.counts_wide, ((u64)S.V + 3) / 4 * 4));
  CK(dalloc(&S.topic_recip, (u64)S.K + 1));
  CK(dalloc(&S.n, VK));
  CK(dalloc(&S.C, S.V));
  CK(dalloc(&S.topic_sum, S.K));
  CK(dalloc(&S.visited, 1));
  CK(dalloc(&S.pois, (u64)kPoisTables * kPoisValues));
  CK(fill_u64(S.visited, 1, 0, S.stream));
  CK(fill_u64(S.topic_sum, S.K, 0, S.stream));
  CK(build_poisson_tables(S.beta, S.pois, S.stream));
  S.host_C.assign(args->C ? args->C : nullptr, args->C ? args->C + S.V : nullptr);
  CK(set_keep_f32((u32)((1u << 20) / ((u64)S.K * sizeof(f32)))));
  if (args->C) {
    HCK(hipMemcpy(S.C, args->C, S.V * sizeof(u32), hipMemcpyHostToDevice));
  } else {
    CK(fill_u32(S.C, S.V, 0, S.stream));
  }
  // initial counts n_vk ~ Pois(floor(C_v / K) + beta) (train.cu:92, polyaurn.cu:61-79)
  dim3 grid((S.K + 255) / 256, (S.V + kPhiRows - 1) / kPhiRows);
  k_init_counts<<<grid, 256, 0, S.stream>>>(S.C, S.n, S.V, S.K, S.beta, (u32)S.seed, (u32)(S.seed >> 32), S.pois,
                                           kPhiRows);
  HCK(hipGetLastError());
  // phi / alias hold the uniform table until the first sample_phi
  CK(fill_u32((u32*)S.phi, VK, 0, S.stream));
  CK(fill_u32((u32*)S.sigma_a, S.V, 0, S.stream));
  CK(fill_u32((u32*)S.counts_fit, ((u64)S.V + 3) / 4, 0, S.stream));  // f32 rows until a Phi step
  CK(fill_u32((u32*)S.counts_wide, ((u64)S.V + 3) / 4, 0, S.stream));
  for (u32 i = 0; i < n_buffers; ++i) {
    Buffer* B = &buffers[i];
    BufferState* bs = new BufferState();
    // non-blocking: the synchronous uploads of another buffer's slice do not wait for this one
    HCK(hipStreamCreateWithFlags(&bs->stream, hipStreamNonBlocking));
    HCK(hipStreamCreateWithFlags(&bs->copy_in, hipStreamNonBlocking));
    HCK(hipStreamCreateWithFlags(&bs->copy_out, hipStreamNonBlocking));
    for (u32 p = 0; p < kMaxParts; ++p) {
      HCK(hipEventCreateWithFlags(&bs->ev_in[p], hipEventDisableTiming));
      HCK(hipEventCreateWithFlags(&bs->ev_swept[p], hipEventDisableTiming));
      HCK(hipEventCreateWithFlags(&bs->ev_out[p], hipEventDisableTiming));
    }
    HCK(hipEventCreateWithFlags(&bs->ev_done, hipEventDisableTiming));
    HCK(hipHostMalloc((void**)&bs->host_z, std::max<u64>(1, args->buffer_size) * sizeof(u16), hipHostMallocDefault));
    CK(dalloc(&bs->heads, kMaxParts));
    B->gpu_z = nullptr;  // z is staged through pinned host memory (u16); the device copy is per slice
    B->gpu_w = nullptr;
    B->gpu_d_len = nullptr;
    B->gpu_d_idx = nullptr;
    B->gpu_K_d = nullptr;
    B->gpu_rng = bs;
    B->stream = &bs->stream;
    bs->owner = B;
    S.buffers.push_back(bs);
  }
  {
    size_t free_b = 0, total_b = 0;
    HCK(hipMemGetInfo(&free_b, &total_b));
    S.cache_auto_budget = free_b / 2;
  }
  HCK(hipStreamSynchronize(S.stream));
  S.ready = true;
  return 0;
}

int do_sample_phi() {
  CK(require_ready());
  HCK(hipDeviceSynchronize());  // buffer sweeps of the previous iteration feed n_wk
  CK(phi_step(S.stream));
  HCK(hipStreamSynchronize(S.stream));
  return 0;
}

int do_sample_z_async(Buffer* B) {
  CK(require_ready());
  if (!B || !B->gpu_rng) return fail("sample_z_async: buffer was not passed to initialize");
  BufferState* bs = (BufferState*)B->gpu_rng;
  if (B->n_docs > S.args->max_D) return fail("buffer holds %u documents, max_D is %u", B->n_docs, S.args->max_D);
  if (B->n_tokens > S.args->buffer_size)
    return fail("buffer holds %u tokens, buffer_size is %u", B->n_tokens, S.args->buffer_size);
  if (bs->pending) return fail("sample_z_async: the previous call on this buffer was not synchronised");
  Segment* sg = nullptr;
  CK(get_segment(bs, B, &sg));
  hipStream_t st = bs->stream;
  CK(fill_u32(bs->heads, kMaxParts, 0, st));
  // Per part p: the host narrows its z into pinned u16 memory and copies it in (copy_in), the
  // part is swept (st), and its topics are copied back to the pinned memory (copy_out);
  // sync_buffer widens each part into the caller's z once its copy has landed.  The host narrows
  // part p + 1 while the device copies and sweeps part p.
  const SweepParams P = sweep_params(sg, S.doc_counter);
  const u32 np = (u32)sg->parts.size();
  if (np > kMaxParts) return fail("slice split into %u parts (limit %u)", np, kMaxParts);
  bs->host_bad = 0;
  for (u32 p = 0; p < np; ++p) {
    const Segment::Part& pt = sg->parts[p];
    const u64 n = pt.tok1 - pt.tok0;
    if (n) {
      bs->host_bad += narrow_topics(B->z + pt.tok0, bs->host_z + pt.tok0, n, S.K);
      HCK(hipMemcpyAsync(sg->z + pt.tok0, bs->host_z + pt.tok0, n * sizeof(u16), hipMemcpyHostToDevice,
                         bs->copy_in));
    }
    HCK(hipEventRecord(bs->ev_in[p], bs->copy_in));
    HCK(hipStreamWaitEvent(st, bs->ev_in[p], 0));
    CK(launch_part_sweep(P, sg, p, bs->heads + p, st));
    HCK(hipEventRecord(bs->ev_swept[p], st));
    HCK(hipStreamWaitEvent(bs->copy_out, bs->ev_swept[p], 0));
    if (n)
      HCK(hipMemcpyAsync(bs->host_z + pt.tok0, sg->z + pt.tok0, n * sizeof(u16), hipMemcpyDeviceToHost,
                         bs->copy_out));
    HCK(hipEventRecord(bs->ev_out[p], bs->copy_out));
  }
  CK(launch_count(sg, S.n, S.K, st));
  HCK(hipEventRecord(bs->ev_done, bs->copy_out));
  HCK(hipStreamWaitEvent(st, bs->ev_done, 0));  // the buffer's stream covers everything
  bs->pending = sg;
  bs->pending_z = B->z;
  S.doc_counter += B->n_docs;
  B->gpu_w = sg->w;
  B->gpu_d_len = sg->doc_off;
  return 0;
}

int do_sync_buffer(Buffer* B) {
  CK(require_ready());
  if (!B || !B->gpu_rng) return fail("sync_buffer: buffer was not passed to initialize");
  BufferState* bs = (BufferState*)B->gpu_rng;
  if (Segment* sg = bs->pending) {
    // widen each part as soon as its copy has landed, while later parts are still on the device
    for (u32 p = 0; p < (u32)sg->parts.size(); ++p) {
      const Segment::Part& pt = sg->parts[p];
      HCK(hipEventSynchronize(bs->ev_out[p]));
      widen_topics(bs->host_z + pt.tok0, bs->pending_z + pt.tok0, pt.tok1 - pt.tok0);
    }
    bs->pending = nullptr;
    bs->pending_z = nullptr;
  }
  HCK(hipStreamSynchronize(bs->stream));
  if (bs->host_bad) return fail("%u topic indicators were >= K = %u (treated as topic 0)", bs->host_bad, S.K);
  return 0;
}

int do_cleanup(Buffer* buffers, u32 n_buffers) {
  if (!S.ready) return fail("cleanup: library not initialised");
  gpulda_serve_close();
  HCK(hipDeviceSynchronize());
  for (u32 i = 0; i < n_buffers; ++i) {
    Buffer* B = &buffers[i];
    BufferState* bs = (BufferState*)B->gpu_rng;
    if (!bs) continue;
    free_segment(bs->seg);
    if (bs->host_z) (void)hipHostFree(bs->host_z);
    dfree(bs->heads);
    for (u32 p = 0; p < kMaxParts; ++p) {
      if (bs->ev_in[p]) (void)hipEventDestroy(bs->ev_in[p]);
      if (bs->ev_swept[p]) (void)hipEventDestroy(bs->ev_swept[p]);
      if (bs->ev_out[p]) (void)hipEventDestroy(bs->ev_out[p]);
    }
    if (bs->ev_done) (void)hipEventDestroy(bs->ev_done);
    if (bs->copy_in) (void)hipStreamDestroy(bs->copy_in);
    if (bs->copy_out) (void)hipStreamDestroy(bs->copy_out);
    if (bs->stream) (void)hipStreamDestroy(bs->stream);
    delete bs;
    B->gpu_z = B->gpu_w = B->gpu_d_len = B->gpu_d_idx = B->gpu_K_d = B->gpu_rng = B->stream = nullptr;
  }
  release_all();
  S.args = nullptr;
  return 0;
}

// ---------------------------------------------------------------------------------------------
// Log-likelihood (double): per-word and per-document partials, summed on the host in order.
// ---------------------------------------------------------------------------------------------
__global__ void k_ll_words(const u32* __restrict__ n, u32 V, u32 K, double beta, double* out_word,
                           unsigned long long* topic_tot) {
  u32 v = blockIdx.x;
  if (v >= V) return;
  __shared__ double part[256];
  double s = 0.0;
  double lgb = lgamma(beta);
  for (u32 k = threadIdx.x; k < K; k += blockDim.x) {
    u32 c = n[(u64)v * K + k];
    if (c) {
      s += lgamma(beta + (double)c) - lgb;
      atomicAdd(&topic_tot[k], (unsigned long long)c);
    }
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (u32 o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out_word[v] = part[0];
}

__global__ void k_ll_docs(const u16* __restrict__ z, const u64* __restrict__ doc_off, u32 n_docs, u32 K, double alpha,
                          double* out_doc) {
  extern __shared__ u32 hist[];
  u32 doc = blockIdx.x;
  if (doc >= n_docs) return;
  __shared__ double part[256];
  for (u32 k = threadIdx.x; k < K; k += blockDim.x) hist[k] = 0;
  __syncthreads();
  u64 b = doc_off[doc];
  u32 L = (u32)(doc_off[doc + 1] - b);
  for (u32 i = threadIdx.x; i < L; i += blockDim.x) atomicAdd(&hist[z[b + i]], 1u);
  __syncthreads();
  double s = 0.0, lga = lgamma(alpha);
  for (u32 k = threadIdx.x; k < K; k += blockDim.x)
    if (hist[k]) s += lgamma(alpha + (double)hist[k]) - lga;
  part[threadIdx.x] = s;
  __syncthreads();
  for (u32 o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out_doc[doc] = part[0] + lgamma((double)K * alpha) - lgamma((double)K * alpha + (double)L);
}

}  // namespace

// =============================================================================================
// The five-call ABI (gpulda_abi.h).
// =============================================================================================
extern "C" void initialize(Args* args, Buffer* buffers, uint32_t n_buffers) {
  abi_die(do_initialize(args, buffers, n_buffers));
}
extern "C" void sample_phi(void) { abi_die(do_sample_phi()); }
extern "C" void sample_z_async(Buffer* buffer) { abi_die(do_sample_z_async(buffer)); }
extern "C" void sync_buffer(Buffer* buffer) { abi_die(do_sync_buffer(buffer)); }
extern "C" void cleanup(Buffer* buffers, uint32_t n_buffers) { abi_die(do_cleanup(buffers, n_buffers)); }

// =============================================================================================
// Extensions (gpulda_ext.h).
// =============================================================================================
extern "C" int gpulda_set_seed(uint64_t seed) {
  g_seed = seed;
  return 0;
}
extern "C" uint64_t gpulda_get_seed(void) { return g_seed; }
extern "C" const char* gpulda_last_error(void) { return g_err.c_str(); }

extern "C" int gpulda_device_info(char* name, int name_len, int* cu_count, uint64_t* hbm_bytes, int* lds_per_block) {
  int dev = 0;
  HCK(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HCK(hipGetDeviceProperties(&p, dev));
  if (name && name_len > 0) {
    snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (uint64_t)p.totalGlobalMem;
  if (lds_per_block) *lds_per_block = max_lds_per_block();
  return 0;
}

extern "C" int gpulda_set_part_tokens(uint64_t min_tokens) {
  if (min_tokens == 0) return fail("part size must be positive");
  g_part_min_tokens = min_tokens;
  return 0;
}

extern "C" int gpulda_set_count_rows(int on) {
  CK(require_ready());
  const bool want = on != 0;
  if (want != S.count_rows) {
    drop_graph();  // a captured iteration holds the Phi step and sweep launches of the old setting
    // rows and flags of an earlier Phi step may be stale: f32 rows until the next Phi step
    if (want) CK(fill_u32((u32*)S.counts_fit, ((u64)S.V + 3) / 4, 0, S.stream));
    S.count_rows = want;
  }
  return 0;
}

extern "C" int gpulda_set_f32_rows(uint32_t n_words) {
  CK(require_ready());
  return set_keep_f32(n_words);
}

extern "C" int gpulda_check_count_rows(int* on, uint64_t* mismatches, uint64_t* words_fit) {
  CK(require_ready());
  if (on) *on = S.count_rows ? 1 : 0;
  if (!S.count_rows) {  // rows and flags of an earlier Phi step may be stale: nothing to check
    if (mismatches) *mismatches = 0;
    if (words_fit) *words_fit = 0;
    return 0;
  }
  HCK(hipStreamSynchronize(S.stream));
  unsigned long long* bad = nullptr;
  CK(dalloc(&bad, 1));
  int rc = fill_u64(bad, 1, 0, 0);
  if (!rc) {
    k_check_counts<<<1024, 256>>>(S.phi, S.phi_counts, S.counts_fit, S.topic_recip, S.V, S.K, bad);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = fail("count-row check failed");
  }
  unsigned long long nb = 0;
  if (!rc && hipMemcpy(&nb, bad, sizeof(nb), hipMemcpyDeviceToHost) != hipSuccess) rc = fail("copy failed");
  dfree(bad);
  CK(rc);
  std::vector<u8> fit(S.V);
  HCK(hipMemcpy(fit.data(), S.counts_fit, S.V, hipMemcpyDeviceToHost));
  u64 nf = 0;
  for (u8 f : fit) nf += f ? 1 : 0;
  if (mismatches) *mismatches = nb;
  if (words_fit) *words_fit = nf;
  return 0;
}

extern "C" int gpulda_set_slice_cache(int mode) {
  if (mode < 0 || mode > 2) return fail("slice cache mode %d (0 off, 1 content-verified, 2 trusted)", mode);
  if (mode != S.slice_cache) {
    // entries are keyed differently in each mode: drop those no buffer has in flight
    for (size_t i = S.cache.size(); i-- > 0;)
      if (!segment_pinned(S.cache[i])) evict_segment(i);
  }
  S.slice_cache = mode;
  return 0;
}

extern "C" int gpulda_set_slice_cache_budget(uint64_t bytes) {
  S.cache_budget = bytes;
  return 0;
}

extern "C" int gpulda_slice_cache_stats(uint32_t* segments, uint64_t* bytes, uint64_t* builds) {
  if (segments) *segments = (uint32_t)S.cache.size();
  if (bytes) *bytes = S.cache_bytes;
  if (builds) *builds = S.cache_builds;
  return 0;
}

extern "C" uint64_t gpulda_slice_hash(const uint32_t* w, const uint32_t* d_len, uint32_t n_docs, uint64_t n_tokens) {
  return slice_content_hash(w, d_len, n_docs, n_tokens);
}

extern "C" uint32_t gpulda_iteration(void) { return S.iter; }

#define COPY_OUT(dst, src, count)                                                                  \
  do {                                                                                             \
    CK(require_ready());                                                                           \
    HCK(hipDeviceSynchronize());                                                                   \
    HCK(hipMemcpy((dst), (src), (u64)(count) * sizeof(*(src)), hipMemcpyDeviceToHost));            \
    return 0;                                                                                      \
  } while (0)

extern "C" int gpulda_copy_phi(float* out) { COPY_OUT(out, S.phi, (u64)S.V * S.K); }
extern "C" int gpulda_copy_sigma_a(float* out) { COPY_OUT(out, S.sigma_a, S.V); }
extern "C" int gpulda_copy_counts(uint32_t* out) { COPY_OUT(out, S.n, (u64)S.V * S.K); }
extern "C" int gpulda_copy_topic_sums(uint64_t* out) {
  COPY_OUT(reinterpret_cast<unsigned long long*>(out), S.topic_sum, S.K);
}

static int split_alias(const AliasSlot* d, u64 count, float* prob, uint32_t* alias) {
  std::vector<AliasSlot> h(count);
  HCK(hipDeviceSynchronize());
  HCK(hipMemcpy(h.data(), d, count * sizeof(AliasSlot), hipMemcpyDeviceToHost));
  for (u64 i = 0; i < count; ++i) {
    if (prob) prob[i] = h[i].prob;
    if (alias) alias[i] = h[i].alias;
  }
  return 0;
}

extern "C" int gpulda_copy_alias(float* prob, uint32_t* alias) {
  CK(require_ready());
  return split_alias(S.alias, (u64)S.V * S.K, prob, alias);
}

extern "C" int gpulda_copy_poisson_tables(float* prob, uint32_t* alias) {
  CK(require_ready());
  return split_alias(S.pois, (u64)kPoisTables * kPoisValues, prob, alias);
}

extern "C" int gpulda_set_counts(const uint32_t* n) {
  CK(require_ready());
  HCK(hipDeviceSynchronize());
  HCK(hipMemcpy(S.n, n, (u64)S.V * S.K * sizeof(u32), hipMemcpyHostToDevice));
  return 0;
}

// Sets the number of Phi steps taken so far; the next sample_phi is step iteration + 1.  Together
// with n_wk restored from a saved z (gpulda_set_counts), a resumed run draws exactly what the
// uninterrupted run would have drawn.
extern "C" int gpulda_set_iteration(uint32_t iteration) {
  CK(require_ready());
  HCK(hipDeviceSynchronize());
  S.iter = iteration;
  CK(fill_u32(S.iter_dev, 1, iteration, S.stream));
  HCK(hipStreamSynchronize(S.stream));
  return 0;
}

extern "C" int gpulda_sweep_stats(uint64_t* visited) {
  COPY_OUT(reinterpret_cast<unsigned long long*>(visited), S.visited, 1);
}

extern "C" int gpulda_reset_stats(void) {
  CK(require_ready());
  CK(fill_u64(S.visited, 1, 0, S.stream));
  HCK(hipStreamSynchronize(S.stream));
  return 0;
}

// ---- resident corpus --------------------------------------------------------------------------
extern "C" int gpulda_corpus_load(const uint32_t* w, const uint32_t* d_len, uint32_t n_docs, const uint32_t* z) {
  CK(require_ready());
  if (!w || !d_len) return fail("corpus_load: w and d_len are required");
  HCK(hipDeviceSynchronize());
  drop_graph();
  free_segment(S.resident);
  S.resident = nullptr;
  u64 N = 0;
  for (u32 i = 0; i < n_docs; ++i) N += d_len[i];
  Segment* sg = new Segment();
  int rc = build_segment(sg, w, d_len, n_docs, N, S.V, S.K);
  if (rc) {
    free_segment(sg);
    return rc;
  }
  if (z) {
    u32* stage = nullptr;
    u32* bad = nullptr;
    CK(dalloc(&stage, N));
    CK(dalloc(&bad, 1));
    CK(fill_u32(bad, 1, 0, S.stream));
    HCK(hipMemcpyAsync(stage, z, N * sizeof(u32), hipMemcpyHostToDevice, S.stream));
    k_z_narrow<<<grid_for(N, 256), 256, 0, S.stream>>>(stage, sg->z, N, S.K, bad);
    HCK(hipGetLastError());
    u32 hbad = 0;
    HCK(hipMemcpyAsync(&hbad, bad, sizeof(u32), hipMemcpyDeviceToHost, S.stream));
    HCK(hipStreamSynchronize(S.stream));
    dfree(stage);
    dfree(bad);
    if (hbad) {
      free_segment(sg);
      return fail("%u initial topic indicators are >= K", hbad);
    }
  } else if (n_docs) {
    k_z_init<<<n_docs, 256, 0, S.stream>>>(sg->z, sg->doc_off, n_docs, 0, S.K, (u32)S.seed, (u32)(S.seed >> 32));
    HCK(hipGetLastError());
    HCK(hipStreamSynchronize(S.stream));
  }
  S.resident = sg;
  return 0;
}

extern "C" int gpulda_corpus_counts_from_z(void) {
  CK(require_ready());
  if (!S.resident) return fail("no resident corpus");
  Segment* sg = S.resident;
  CK(fill_u32(S.n, (u64)S.V * S.K, 0, S.stream));
  CK(launch_count(sg, S.n, S.K, S.stream));
  HCK(hipStreamSynchronize(S.stream));
  return 0;
}

extern "C" int gpulda_corpus_copy_z(uint32_t* z) {
  CK(require_ready());
  if (!S.resident) return fail("no resident corpus");
  Segment* sg = S.resident;
  u64 N = sg->n_tokens;
  u32* stage = nullptr;
  CK(dalloc(&stage, N));
  if (N) {
    k_z_widen<<<grid_for(N, 256), 256, 0, S.stream>>>(sg->z, stage, N);
    HCK(hipGetLastError());
    HCK(hipMemcpyAsync(z, stage, N * sizeof(u32), hipMemcpyDeviceToHost, S.stream));
  }
  HCK(hipStreamSynchronize(S.stream));
  dfree(stage);
  return 0;
}

namespace {
// One full iteration of the resident corpus on stream st; ev (5 events or null) brackets the
// Poisson draw, alias rows (+ n_wk clear), sweep and reduction.
int resident_iteration(Segment* sg, hipStream_t st, hipEvent_t* ev) {
  if (ev) HCK(hipEventRecord(ev[0], st));
  CK(phi_step(st, ev ? ev[1] : nullptr));
  if (ev) HCK(hipEventRecord(ev[2], st));
  CK(launch_sweep_schedule(sweep_params(sg, 0), sg, st));
  if (ev) HCK(hipEventRecord(ev[3], st));
  CK(launch_count(sg, S.n, S.K, st));
  if (ev) HCK(hipEventRecord(ev[4], st));
  S.doc_counter = sg->n_docs;
  return 0;
}

// Captures one iteration into a graph (the iteration number lives in device memory, so a replay
// is the next iteration).
int capture_graph(Segment* sg) {
  drop_graph();
  hipGraph_t g = nullptr;
  HCK(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
  const u32 iter_before = S.iter;
  int rc = resident_iteration(sg, S.stream, nullptr);
  hipError_t e = hipStreamEndCapture(S.stream, &g);
  S.iter = iter_before;  // capture does not run anything
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HCK(e);
  hipError_t ei = hipGraphInstantiate(&S.graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HCK(ei);
  return 0;
}
}  // namespace

extern "C" int gpulda_corpus_iterate(uint32_t n_iters, float* stage_ms) {
  CK(require_ready());
  if (!S.resident) return fail("no resident corpus");
  Segment* sg = S.resident;
  hipStream_t st = S.stream;
  double acc[4] = {0, 0, 0, 0};
  if (S.use_graph) {
    if (!S.graph) CK(capture_graph(sg));
    HCK(hipEventRecord(S.ev[0], st));
    for (u32 it = 0; it < n_iters; ++it) {
      HCK(hipGraphLaunch(S.graph, st));
      S.iter += 1;
    }
    HCK(hipEventRecord(S.ev[4], st));
    HCK(hipEventSynchronize(S.ev[4]));
    float ms = 0;
    HCK(hipEventElapsedTime(&ms, S.ev[0], S.ev[4]));
    if (stage_ms) {  // a graph replays the whole iteration: only the total is timed
      stage_ms[0] = stage_ms[1] = stage_ms[3] = 0.0f;
      stage_ms[2] = n_iters ? ms / n_iters : 0.0f;
    }
    S.doc_counter = sg->n_docs;
    return 0;
  }
  for (u32 it = 0; it < n_iters; ++it) {
    CK(resident_iteration(sg, st, stage_ms ? S.ev : nullptr));
    if (stage_ms) {
      HCK(hipEventSynchronize(S.ev[4]));
      for (int k = 0; k < 4; ++k) {
        float ms = 0;
        HCK(hipEventElapsedTime(&ms, S.ev[k], S.ev[k + 1]));
        acc[k] += ms;
      }
    }
  }
  HCK(hipStreamSynchronize(st));
  if (stage_ms)
    for (int k = 0; k < 4; ++k) stage_ms[k] = n_iters ? (float)(acc[k] / n_iters) : 0.0f;
  return 0;
}

extern "C" int gpulda_corpus_schedule(int schedule, uint32_t waves_per_cu, uint32_t block_min_len) {
  CK(require_ready());
  if (schedule < kSchedBuckets || schedule > kSchedPersistentSplit) return fail("unknown schedule %d", schedule);
  S.schedule = schedule;
  S.waves_per_cu = waves_per_cu ? waves_per_cu : 16;
  S.block_min_len = block_min_len;
  drop_graph();
  return 0;
}

extern "C" int gpulda_corpus_use_graph(int enabled) {
  CK(require_ready());
  S.use_graph = enabled != 0;
  drop_graph();
  return 0;
}

// Times one sweep of the resident corpus under each candidate schedule (z is saved before and
// restored after every trial, so the chain does not move), then applies the fastest.
// table: 4 floats per candidate {schedule, waves_per_cu, block_min_len, ms}.
extern "C" int gpulda_corpus_autotune(float* table, uint32_t max_candidates, uint32_t* n_candidates, int* chosen) {
  CK(require_ready());
  if (!S.resident) return fail("no resident corpus");
  if (S.iter == 0) return fail("autotune needs a Phi step first (iterate once)");
  Segment* sg = S.resident;
  struct Cand {
    int sched;
    u32 waves;
    u32 minlen;
    u32 rows;  // 1: count rows where the persistent sweep can read them (as the last Phi step wrote them)
  };
  // count rows only as the last Phi step left them: when they are on, each persistent candidate
  // is also timed on f32 rows, and switching them off applies from the next Phi step
  const u32 cr = S.count_rows ? 1u : 0u;
  std::vector<Cand> cands = {{kSchedBuckets, 0, 0, 0}};
  for (u32 r = 0; r <= cr; ++r) {
    for (u32 w : {4u, 8u, 16u, 32u}) cands.push_back({kSchedPersistent, w, 0, r});
    const u32 longest = sg->lens_desc.empty() ? 0 : sg->lens_desc[0];
    for (u32 m : {2048u, 8192u})
      if (longest >= m) cands.push_back({kSchedPersistentBlock, 16, m, r});
    for (u32 m : {512u, 1024u, 2048u})
      if (longest > m) cands.push_back({kSchedPersistentSplit, 32, m, r});
  }
  u16* saved = nullptr;
  CK(dalloc(&saved, sg->n_tokens));
  unsigned long long vis0 = 0;
  HCK(hipStreamSynchronize(S.stream));
  HCK(hipMemcpy(&vis0, S.visited, 8, hipMemcpyDeviceToHost));
  HCK(hipMemcpyAsync(saved, sg->z, sg->n_tokens * 2, hipMemcpyDeviceToDevice, S.stream));
  float best = 1e30f;
  int best_i = 0;
  for (size_t i = 0; i < cands.size(); ++i) {
    S.schedule = cands[i].sched;
    S.waves_per_cu = cands[i].waves ? cands[i].waves : 16;
    S.block_min_len = cands[i].minlen;
    SweepParams P = sweep_params(sg, 0);
    if (!cands[i].rows) {
      P.phi_counts = nullptr;
      P.counts_fit = nullptr;
      P.topic_recip = nullptr;
    }
    HCK(hipEventRecord(S.ev[0], S.stream));
    int rc = launch_sweep_schedule(P, sg, S.stream);
    HCK(hipEventRecord(S.ev[1], S.stream));
    HCK(hipMemcpyAsync(sg->z, saved, sg->n_tokens * 2, hipMemcpyDeviceToDevice, S.stream));
    HCK(hipStreamSynchronize(S.stream));
    if (rc) {
      dfree(saved);
      return rc;
    }
    float ms = 0;
    HCK(hipEventElapsedTime(&ms, S.ev[0], S.ev[1]));
    if (table && i < max_candidates) {
      table[5 * i + 0] = (float)cands[i].sched;
      table[5 * i + 1] = (float)cands[i].waves;
      table[5 * i + 2] = (float)cands[i].minlen;
      table[5 * i + 3] = (float)cands[i].rows;
      table[5 * i + 4] = ms;
    }
    if (ms < best) {
      best = ms;
      best_i = (int)i;
    }
  }
  dfree(saved);
  HCK(hipMemcpy(S.visited, &vis0, 8, hipMemcpyHostToDevice));
  S.schedule = cands[best_i].sched;
  S.waves_per_cu = cands[best_i].waves ? cands[best_i].waves : 16;
  S.block_min_len = cands[best_i].minlen;
  if (cr && !cands[best_i].rows) S.count_rows = false;  // the Phi step stops writing them
  drop_graph();
  if (n_candidates) *n_candidates = (u32)cands.size();
  if (chosen) *chosen = best_i;
  return 0;
}

extern "C" int gpulda_corpus_loglik(double* out) {
  CK(require_ready());
  if (!S.resident) return fail("no resident corpus");
  Segment* sg = S.resident;
  double* dw = nullptr;
  double* dd = nullptr;
  unsigned long long* tot = nullptr;
  CK(dalloc(&dw, S.V));
  CK(dalloc(&dd, std::max<u32>(1, sg->n_docs)));
  CK(dalloc(&tot, S.K));
  CK(fill_u64(tot, S.K, 0, S.stream));
  k_ll_words<<<S.V, 256, 0, S.stream>>>(S.n, S.V, S.K, (double)S.beta, dw, tot);
  HCK(hipGetLastError());
  if (sg->n_docs) {
    k_ll_docs<<<sg->n_docs, 256, (u64)S.K * sizeof(u32), S.stream>>>(sg->z, sg->doc_off, sg->n_docs, S.K,
                                                                       (double)S.alpha, dd);
    HCK(hipGetLastError());
  }
  std::vector<double> hw(S.V), hd(sg->n_docs);
  std::vector<unsigned long long> ht(S.K);
  HCK(hipStreamSynchronize(S.stream));
  HCK(hipMemcpy(hw.data(), dw, S.V * sizeof(double), hipMemcpyDeviceToHost));
  if (sg->n_docs) HCK(hipMemcpy(hd.data(), dd, sg->n_docs * sizeof(double), hipMemcpyDeviceToHost));
  HCK(hipMemcpy(ht.data(), tot, S.K * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  dfree(dw);
  dfree(dd);
  dfree(tot);
  double ll = 0.0;
  for (double x : hw) ll += x;
  double vb = (double)S.V * (double)S.beta;
  for (u32 k = 0; k < S.K; ++k) ll += std::lgamma(vb) - std::lgamma(vb + (double)ht[k]);
  for (double x : hd) ll += x;
  *out = ll;
  return 0;
}

// Plain copy kernel of known size, for scaling hardware byte counters (FETCH_SIZE/WRITE_SIZE).
__global__ void k_calibrate_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, u64 n) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) dst[i] = src[i];
}

extern "C" int gpulda_calibrate_copy(uint64_t bytes, uint32_t reps, float* ms_per_copy) {
  const u64 n = bytes / sizeof(uint4);
  uint4 *a = nullptr, *b = nullptr;
  CK(dalloc(&a, n));
  CK(dalloc(&b, n));
  hipEvent_t e0, e1;
  HCK(hipEventCreate(&e0));
  HCK(hipEventCreate(&e1));
  k_fill_u32<<<grid_for(n * 4, 256), 256>>>((u32*)a, n * 4, 1u);
  HCK(hipEventRecord(e0, 0));
  for (u32 r = 0; r < reps; ++r) k_calibrate_copy<<<4096, 256>>>(a, b, n);
  HCK(hipEventRecord(e1, 0));
  HCK(hipEventSynchronize(e1));
  float ms = 0;
  HCK(hipEventElapsedTime(&ms, e0, e1));
  if (ms_per_copy) *ms_per_copy = reps ? ms / reps : 0.0f;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dfree(a);
  dfree(b);
  return 0;
}

extern "C" int gpulda_corpus_release(void) {
  HCK(hipDeviceSynchronize());
  drop_graph();
  free_segment(S.resident);
  S.resident = nullptr;
  return 0;
}

// ---- fold-in server ----------------------------------------------------------------------------
// Continuous batching against the resident model (phi, sigma_a and alias tables stay in HBM):
//   submit   admits a request if a slot and a span of the token arena (bounded by the HBM budget)
//            are free and stages its words in host-visible memory; otherwise the caller retries;
//   flush    publishes every admitted, not yet published request to the request ring;
//   poll     reports requests whose completion flag the device raised;
//   result   returns a finished request's topics and frees its slot and span.
// A pool of worker wavefronts (k_serve_worker, waves_per_cu per CU) stays resident and takes
// requests from the ring as soon as they are published, one wavefront per request, so requests
// that arrive while others run start at once and overlap on the device.  A pool stops claiming
// after kPoolLifetime seconds and its kernel ends once its last request is done; the host starts a
// younger pool on another of kPools streams once the newest is half way through (and whenever none
// is running while requests are outstanding).  An expiring pool may take as long as its longest
// request to drain, so a third stream keeps one pool claiming meanwhile.
// A request's randomness is keyed by its request id, so its result is the same whatever runs
// beside it.
namespace {
constexpr double kPoolLifetime = 2.0;  // seconds a pool keeps claiming requests
constexpr int kPools = 3;              // pool streams (with the feeder's: four hardware queues)

double g_serve_deadline = 30.0;  // seconds from publication to a lost report (gpulda_serve_set_deadline)

struct ServePool {
  hipStream_t st = nullptr;
  double launched = 0.0;  // host seconds
  bool active = false;
};

struct ServeState {
  bool open = false;
  u32 sweeps = 50;
  u32 n_slots = 0;
  u32 waves_per_cu = 2;
  u64 arena = 0;            // tokens
  u32* d_w = nullptr;       // arena word ids (copied in by DMA at each flush)
  u16* d_z = nullptr;       // arena topics
  u32* d_rec = nullptr;     // per slot request record (kServeRecWords, one line each)
  u32* d_ring = nullptr;    // request ring (n_slots entries)
  u32* d_ctl = nullptr;     // [0] tail, [1] head, [2] shutdown
  u32* d_meta = nullptr;    // records as the host sends them (2 * n_slots, a ring)
  u32* h_w = nullptr;       // pinned staging of the arena word ids
  u16* h_z = nullptr;       // host-visible final topics (arena offsets), written by the device
  u64* h_off = nullptr;     // slot -> span offset
  u32* h_len = nullptr;     // slot -> span length
  u32* h_rid = nullptr;     // slot -> request id
  u32* h_done = nullptr;    // host-visible slot -> completion flag (the device's fingerprint)
  u32* h_drec = nullptr;    // host-visible slot -> what the serving wave read and wrote
  u32* h_meta = nullptr;    // pinned request records of the flushes (2 * n_slots records, a ring)
  u32 meta_head = 0;        // next free record
  std::vector<u32> expect;  // slot -> fingerprint of the staged request
  std::vector<u32> wsum;    // slot -> word sum of the staged request
  std::vector<u32> pub_at;  // slot -> ring index it was published at
  u32 tail = 0;
  u64 lifetime_ticks = 0;
  std::vector<u32> free_slots;
  std::vector<std::pair<u64, u64>> free_spans;  // (offset, length), sorted by offset
  std::vector<u32> pending;   // admitted, not published
  ServeBook book;             // published requests: reported, lost, quarantined (serve_book.h)
  ServePool pools[kPools];
  hipStream_t feed = nullptr;
  u64 lds = 0;
};
ServeState SV;

double host_seconds() {
  timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

// Spans start on 32-token (128-byte) boundaries, so no two requests share a cache line of words.
u64 serve_span_tokens(u32 L) { return ((u64)L + 31) & ~(u64)31; }

int serve_alloc_span(u64 len, u64* off) {
  for (size_t i = 0; i < SV.free_spans.size(); ++i) {
    if (SV.free_spans[i].second >= len) {
      *off = SV.free_spans[i].first;
      SV.free_spans[i].first += len;
      SV.free_spans[i].second -= len;
      if (SV.free_spans[i].second == 0) SV.free_spans.erase(SV.free_spans.begin() + (long)i);
      return 0;
    }
  }
  return 1;
}

void serve_free_span(u64 off, u64 len) {
  if (len == 0) return;
  auto it = std::lower_bound(SV.free_spans.begin(), SV.free_spans.end(), std::make_pair(off, (u64)0));
  it = SV.free_spans.insert(it, {off, len});
  // coalesce with the neighbours
  size_t i = (size_t)(it - SV.free_spans.begin());
  if (i + 1 < SV.free_spans.size() && SV.free_spans[i].first + SV.free_spans[i].second == SV.free_spans[i + 1].first) {
    SV.free_spans[i].second += SV.free_spans[i + 1].second;
    SV.free_spans.erase(SV.free_spans.begin() + (long)i + 1);
  }
  if (i > 0 && SV.free_spans[i - 1].first + SV.free_spans[i - 1].second == SV.free_spans[i].first) {
    SV.free_spans[i - 1].second += SV.free_spans[i].second;
    SV.free_spans.erase(SV.free_spans.begin() + (long)i);
  }
}

// Pinned host memory mapped into the device address space.  Kernels read it only at their start
// (a launch makes the host's earlier stores visible); the device writes it only with system-scope
// release stores, which reach host memory before the store that follows them.
template <class T>
int host_alloc(T** p, u64 n) {
  *p = nullptr;
  HCK(hipHostMalloc((void**)p, std::max<u64>(n, 1) * sizeof(T), hipHostMallocMapped));
  return 0;
}
template <class T>
T* dev_view(T* h) {
  void* d = nullptr;
  if (hipHostGetDevicePointer(&d, (void*)h, 0) != hipSuccess) return nullptr;
  return (T*)d;
}

void serve_release() {
  if (SV.d_ctl && SV.feed) {
    k_fill_u32<<<1, 64, 0, SV.feed>>>(SV.d_ctl + 2, 1, 1u);  // shutdown: every pool wave leaves
    (void)hipStreamSynchronize(SV.feed);
  }
  for (ServePool& pl : SV.pools)
    if (pl.st) {
      (void)hipStreamSynchronize(pl.st);
      (void)hipStreamDestroy(pl.st);
    }
  if (SV.feed) (void)hipStreamDestroy(SV.feed);
  for (void* d : {(void*)SV.d_w, (void*)SV.d_z, (void*)SV.d_rec, (void*)SV.d_ring, (void*)SV.d_ctl, (void*)SV.d_meta})
    if (d) (void)hipFree(d);
  for (void* h : {(void*)SV.h_w, (void*)SV.h_z, (void*)SV.h_off, (void*)SV.h_len, (void*)SV.h_rid, (void*)SV.h_done,
                  (void*)SV.h_meta, (void*)SV.h_drec})
    if (h) (void)hipHostFree(h);
  SV = ServeState();
}

ServeRing serve_ring() {
  ServeRing R;
  R.ring = SV.d_ring;
  R.ring_size = SV.n_slots;
  R.tail = SV.d_ctl;
  R.head = SV.d_ctl + 1;
  R.shutdown = SV.d_ctl + 2;
  R.lifetime_ticks = SV.lifetime_ticks;
  R.rec = SV.d_rec;
  R.arena = SV.arena;
  R.V = S.V;
  R.done_rec = dev_view(SV.h_drec);
  return R;
}

SweepParams serve_params() {
  SweepParams P;
  P.w = SV.d_w;
  P.z = SV.d_z;
  P.doc_off = nullptr;
  P.doc_order = nullptr;
  P.n_docs = SV.n_slots;
  P.doc_base = 0;
  P.iter = 0;
  P.iter_ptr = nullptr;
  P.k0 = (u32)S.seed;
  P.k1 = (u32)(S.seed >> 32);
  P.K = S.K;
  P.phi = S.phi;
  P.sigma_a = S.sigma_a;
  P.alias = S.alias;
  P.visited = nullptr;
  return P;
}

int launch_pool(ServePool& pl) {
  FoldParams F;
  F.slots = nullptr;
  F.n_slots = SV.n_slots;
  F.span_off = nullptr;  // the worker reads the request records instead
  F.span_len = nullptr;
  F.req_id = nullptr;
  F.sweeps = SV.sweeps;
  F.z_out = dev_view(SV.h_z);
  F.done = dev_view(SV.h_done);
  F.done_value = 0;
  if (!F.done || !F.z_out || !dev_view(SV.h_drec)) return fail("host-mapped result buffers unavailable");
  const u32 waves = std::max<u32>(1, (u32)S.cu_count * SV.waves_per_cu);
  k_serve_worker<<<waves, 64, SV.lds, pl.st>>>(serve_params(), F, serve_ring());
  HCK(hipGetLastError());
  pl.launched = host_seconds();
  pl.active = true;
  return 0;
}

// Keeps a claiming pool on the device while requests are outstanding.
// Lost requests the caller has collected, released while no pool kernel runs (no worker can touch
// them): the ring entry (if it still names the slot) and the record are cleared on the device,
// then the slot and its span are free again.
int serve_release_lost() {
  const std::vector<u32> rel = SV.book.release(true);
  if (rel.empty()) return 0;
  HCK(hipStreamSynchronize(SV.feed));  // no publication in flight
  const u32 zero = 0;
  for (u32 s : rel) {
    const u32 pos = SV.pub_at[s] % SV.n_slots;
    u32 e = 0;
    HCK(hipMemcpy(&e, SV.d_ring + pos, sizeof(u32), hipMemcpyDeviceToHost));
    if (e == s + 1u) HCK(hipMemcpy(SV.d_ring + pos, &zero, sizeof(u32), hipMemcpyHostToDevice));
    HCK(hipMemcpy(SV.d_rec + (u64)kServeRecWords * s, &zero, sizeof(u32), hipMemcpyHostToDevice));
    __atomic_store_n(&SV.h_done[s], 0u, __ATOMIC_RELEASE);
    serve_free_span(SV.h_off[s], serve_span_tokens(SV.h_len[s]));
    SV.free_slots.push_back(s);
  }
  return 0;
}

int serve_keep_pool() {
  const double now = host_seconds();
  double newest = -1.0;
  for (ServePool& pl : SV.pools) {
    if (pl.active && hipStreamQuery(pl.st) == hipSuccess) pl.active = false;
    if (pl.active) newest = std::max(newest, pl.launched);
  }
  if (newest < 0.0 && !SV.book.quarantine.empty()) CK(serve_release_lost());
  if (SV.book.inflight.empty()) return 0;
  const bool none = newest < 0.0;
  const bool ageing = !none && now - newest > 0.5 * kPoolLifetime;
  if (!none && !ageing) return 0;
  for (ServePool& pl : SV.pools)
    if (!pl.active) return launch_pool(pl);
  return 0;  // both pools still running: the younger one is claiming
}
}  // namespace

extern "C" int gpulda_serve_open(uint64_t hbm_budget_bytes, uint32_t sweeps, uint32_t waves_per_cu, uint32_t max_slots) {
  CK(require_ready());
  if (S.iter == 0) return fail("serve_open needs a trained model (run at least one Phi step)");
  if (SV.open) serve_release();
  SV.sweeps = sweeps;
  SV.n_slots = max_slots ? max_slots : 65536;
  SV.waves_per_cu = std::max<u32>(1, waves_per_cu);
  SV.arena = std::max<u64>(hbm_budget_bytes / (sizeof(u32) + sizeof(u16)), 1);  // words and topics per token
  SV.lds = doc_state_bytes(S.K, S.K);
  if (SV.lds > (u64)max_lds_per_block()) return fail("doc-topic map exceeds LDS");
  int dev = 0, rate_khz = 0;
  HCK(hipGetDevice(&dev));
  HCK(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
  if (rate_khz <= 0) return fail("device wall clock rate unavailable");
  SV.lifetime_ticks = (u64)(kPoolLifetime * 1000.0 * (double)rate_khz);
  int rc = 0;
  if ((rc = dalloc(&SV.d_w, SV.arena)) || (rc = dalloc(&SV.d_z, SV.arena)) ||
      (rc = dalloc(&SV.d_rec, (u64)kServeRecWords * SV.n_slots)) || (rc = dalloc(&SV.d_ring, SV.n_slots)) ||
      (rc = dalloc(&SV.d_ctl, 4)) || (rc = dalloc(&SV.d_meta, 2ull * kServeMetaWords * SV.n_slots)) ||
      (rc = host_alloc(&SV.h_w, SV.arena)) ||
      (rc = host_alloc(&SV.h_z, SV.arena)) || (rc = host_alloc(&SV.h_off, SV.n_slots)) ||
      (rc = host_alloc(&SV.h_len, SV.n_slots)) || (rc = host_alloc(&SV.h_rid, SV.n_slots)) ||
      (rc = host_alloc(&SV.h_done, SV.n_slots)) || (rc = host_alloc(&SV.h_meta, 2ull * kServeMetaWords * SV.n_slots)) ||
      (rc = host_alloc(&SV.h_drec, (u64)kServeRecWords * SV.n_slots))) {
    serve_release();
    return rc;
  }
  HCK(hipMemset(SV.d_ring, 0, SV.n_slots * sizeof(u32)));
  HCK(hipMemset(SV.d_rec, 0, (u64)kServeRecWords * SV.n_slots * sizeof(u32)));
  HCK(hipMemset(SV.d_ctl, 0, 4 * sizeof(u32)));
  for (ServePool& pl : SV.pools) HCK(hipStreamCreateWithFlags(&pl.st, hipStreamNonBlocking));
  HCK(hipStreamCreateWithFlags(&SV.feed, hipStreamNonBlocking));
  HCK(hipDeviceSynchronize());
  SV.free_slots.resize(SV.n_slots);
  for (u32 i = 0; i < SV.n_slots; ++i) SV.free_slots[i] = SV.n_slots - 1 - i;
  SV.expect.assign(SV.n_slots, 0);
  SV.wsum.assign(SV.n_slots, 0);
  SV.pub_at.assign(SV.n_slots, 0);
  SV.book.reset(SV.n_slots);
  SV.book.deadline = g_serve_deadline;
  SV.free_spans = {{0, SV.arena}};
  memset(SV.h_done, 0, SV.n_slots * sizeof(u32));
  memset(SV.h_drec, 0, (u64)kServeRecWords * SV.n_slots * sizeof(u32));
  SV.open = true;
  return 0;
}

// Returns 0 when admitted (*slot set), 1 when the request must wait (no slot or arena space).
extern "C" int gpulda_serve_submit(const uint32_t* w, uint32_t L, uint32_t request_id, uint32_t* slot) {
  if (!SV.open) return fail("server not open");
  if (L == 0 || L > kMaxDocLen) return fail("request length %u outside 1..%u", L, kMaxDocLen);
  for (u32 i = 0; i < L; ++i)
    if (w[i] >= S.V) return fail("request word id %u >= V", w[i]);
  if (SV.free_slots.empty()) return 1;
  u64 off = 0;
  if (serve_alloc_span(serve_span_tokens(L), &off)) return 1;
  u32 s = SV.free_slots.back();
  SV.free_slots.pop_back();
  memcpy(SV.h_w + off, w, (u64)L * sizeof(u32));
  u32 word_sum = 0;
  for (u32 i = 0; i < L; ++i) word_sum += w[i] * (2u * i + 1u);
  SV.expect[s] = serve_fingerprint(word_sum, L, request_id);
  SV.wsum[s] = word_sum;
  SV.h_off[s] = off;
  SV.h_len[s] = L;
  SV.h_rid[s] = request_id;
  __atomic_store_n(&SV.h_done[s], 0u, __ATOMIC_RELEASE);
  SV.pending.push_back(s);
  *slot = s;
  return 0;
}

extern "C" int gpulda_serve_flush(void) {
  if (!SV.open) return fail("server not open");
  if (!SV.pending.empty()) {
    const u32 n = (u32)SV.pending.size();
    // records live in a ring of 2 * n_slots; before it wraps onto records whose copies may still
    // be pending, wait for the feeder stream
    if (SV.meta_head + n > 2 * SV.n_slots) {
      HCK(hipStreamSynchronize(SV.feed));
      SV.meta_head = 0;
    }
    u32* rec = SV.h_meta + (u64)SV.meta_head * kServeMetaWords;
    for (u32 i = 0; i < n; ++i) {
      const u32 sl = SV.pending[i];
      const u64 off = SV.h_off[sl];
      const u32 L = SV.h_len[sl];
      rec[i * kServeMetaWords + 0] = sl;
      rec[i * kServeMetaWords + 1] = L;
      rec[i * kServeMetaWords + 2] = SV.h_rid[sl];
      rec[i * kServeMetaWords + 3] = (u32)off;
      rec[i * kServeMetaWords + 4] = (u32)(off >> 32);
      rec[i * kServeMetaWords + 5] = SV.wsum[sl];
      SV.pub_at[sl] = SV.tail + i;
      HCK(hipMemcpyAsync(SV.d_w + off, SV.h_w + off, (u64)L * sizeof(u32), hipMemcpyHostToDevice, SV.feed));
    }
    u32* drec = SV.d_meta + (u64)SV.meta_head * kServeMetaWords;
    HCK(hipMemcpyAsync(drec, rec, (u64)n * kServeMetaWords * sizeof(u32), hipMemcpyHostToDevice, SV.feed));
    ServeFeed Fd;
    Fd.meta = drec;
    Fd.base = SV.tail;
    Fd.count = n;
    // at most n_slots requests are admitted at a time, so a ring entry is taken before the ring
    // wraps onto it
    k_serve_feed<<<(n + 63) / 64, 64, 0, SV.feed>>>(Fd, serve_ring());
    HCK(hipGetLastError());
    SV.meta_head += n;
    SV.tail += n;
    const double now = host_seconds();
    for (u32 sl : SV.pending) SV.book.published(sl, now);
    SV.pending.clear();
  }
  return serve_keep_pool();
}

// Slots whose fold-in finished since the last poll (at most max).
extern "C" int gpulda_serve_poll(uint32_t* done_slots, uint32_t max, uint32_t* n_done) {
  if (!SV.open) return fail("server not open");
  // finished slots, and slots whose request got no result within the deadline (collected as
  // failed by gpulda_serve_result)
  *n_done = SV.book.poll([](u32 sl) { return __atomic_load_n(&SV.h_done[sl], __ATOMIC_ACQUIRE) != 0u; },
                         host_seconds(), done_slots, max);
  return serve_keep_pool();
}

// Copies a finished request's topics (u32) and frees its slot.
extern "C" int gpulda_serve_result(uint32_t slot, uint32_t* z, uint32_t* request_id) {
  if (!SV.open) return fail("server not open");
  if (slot < SV.n_slots && SV.book.lost[slot]) {
    if (request_id) *request_id = SV.h_rid[slot];
    SV.book.collected_lost(slot);  // slot and span come back once no worker runs
    return fail("slot %u: request %u got no result within %.1f s of its publication", slot, SV.h_rid[slot],
                SV.book.deadline);
  }
  if (slot >= SV.n_slots || !__atomic_load_n(&SV.h_done[slot], __ATOMIC_ACQUIRE))
    return fail("slot %u has no finished request", slot);
  const u64 off = SV.h_off[slot];
  const u32 L = SV.h_len[slot];
  const u32 flag = __atomic_load_n(&SV.h_done[slot], __ATOMIC_ACQUIRE);
  const u32* dr = SV.h_drec + (u64)kServeRecWords * slot;
  int rc = 0;
  if (flag != SV.expect[slot]) {
    rc = fail("slot %u: the device processed different request data (fingerprint %08x, staged %08x); it read ring "
              "index %d, L %u, id %u, offset %u, word sum %08x; the host published ring index %u with L %u, id %u, "
              "offset %u, word sum %08x",
              slot, flag, SV.expect[slot], (int)dr[0] - 1, dr[1], dr[2], dr[3], dr[5], SV.pub_at[slot], L,
              SV.h_rid[slot], (u32)off, SV.wsum[slot]);
  } else {
    // the topics arrive by system-scope stores ordered before the flag; their checksum is in the
    // record, and the copy is read again (briefly) until it matches
    const double t0 = host_seconds();
    for (;;) {
      u32 zs = 0;
      for (u32 i = 0; i < L; ++i) zs += ((u32)__atomic_load_n(&SV.h_z[off + i], __ATOMIC_ACQUIRE) + 1u) * (2u * i + 1u);
      if (zs == __atomic_load_n(&dr[6], __ATOMIC_ACQUIRE)) break;
      if (host_seconds() - t0 > 1.0) {
        rc = fail("slot %u: the topics did not arrive intact", slot);
        break;
      }
    }
    if (rc == 0 && z)
      for (u32 i = 0; i < L; ++i) z[i] = SV.h_z[off + i];
  }
  if (request_id) *request_id = SV.h_rid[slot];
  __atomic_store_n(&SV.h_done[slot], 0u, __ATOMIC_RELEASE);
  serve_free_span(off, serve_span_tokens(L));
  SV.free_slots.push_back(slot);
  return rc;
}

extern "C" int gpulda_serve_set_deadline(double seconds) {
  if (!(seconds > 0.0)) return fail("serve deadline must be positive");
  g_serve_deadline = seconds;
  if (SV.open) SV.book.deadline = seconds;
  return 0;
}

extern "C" int gpulda_serve_stats(uint32_t* pending, uint32_t* inflight, uint64_t* free_tokens) {
  if (!SV.open) return fail("server not open");
  if (pending) *pending = (u32)SV.pending.size();
  if (inflight) *inflight = (u32)SV.book.inflight.size();
  if (free_tokens) {
    u64 f = 0;
    for (auto& sp : SV.free_spans) f += sp.second;
    *free_tokens = f;
  }
  return 0;
}

extern "C" int gpulda_serve_close(void) {
  if (SV.open) serve_release();
  return 0;
}

// ---- stage entry points -----------------------------------------------------------------------
namespace {
// RAII bundle of temporary device buffers for the stage entry points.
struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    (void)hipDeviceSynchronize();
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <class T>
  int alloc(T** p, u64 n) {
    CK(dalloc(p, n));
    ptrs.push_back((void*)*p);
    return 0;
  }
  template <class T>
  int upload(T** p, const T* h, u64 n) {
    CK(alloc(p, n));
    if (n) HCK(hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice));
    return 0;
  }
};

int pack_alias(Scratch& sc, const float* prob, const uint32_t* alias, u64 n, AliasSlot** out) {
  std::vector<AliasSlot> h(n);
  for (u64 i = 0; i < n; ++i) {
    h[i].prob = prob[i];
    h[i].alias = alias[i];
  }
  return sc.upload(out, h.data(), n);
}

int stage_tables(Scratch& sc, f32 beta, AliasSlot** out) {
  CK(sc.alloc(out, (u64)kPoisTables * kPoisValues));
  return build_poisson_tables(beta, *out, 0);
}
}  // namespace

extern "C" int gpulda_stage_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint32_t n) {
  Scratch sc;
  u32 *dc, *dk, *dout;
  CK(sc.upload(&dc, ctr, 4ull * n));
  CK(sc.upload(&dk, key, 2ull * n));
  CK(sc.alloc(&dout, 4ull * n));
  if (n) {
    k_stage_philox<<<(n + 255) / 256, 256>>>(dc, dk, dout, n);
    HCK(hipGetLastError());
    HCK(hipMemcpy(out, dout, 16ull * n, hipMemcpyDeviceToHost));
  }
  return 0;
}

extern "C" int gpulda_stage_scan64(const float* in, float* out, uint32_t n_groups) {
  Scratch sc;
  f32 *di, *dout;
  CK(sc.upload(&di, in, 64ull * n_groups));
  CK(sc.alloc(&dout, 64ull * n_groups));
  if (n_groups) {
    k_stage_scan<<<n_groups, 64>>>(di, dout, n_groups);
    HCK(hipGetLastError());
    HCK(hipMemcpy(out, dout, 256ull * n_groups, hipMemcpyDeviceToHost));
  }
  return 0;
}

extern "C" int gpulda_stage_map(const uint32_t* ops, uint32_t n_ops, uint32_t cap, uint32_t K, uint32_t* table_out) {
  if (cap == 0 || cap % kMapLine) return fail("map capacity must be a positive multiple of 64");
  if (K == 0 || K > kMaxSweepTopics) return fail("K = %u out of range (1..%u)", K, kMaxSweepTopics);
  if ((u64)doc_counts_bytes(K) + (u64)cap * 2 > (u64)max_lds_per_block()) return fail("map capacity exceeds LDS");
  for (u32 i = 0; i < n_ops; ++i)
    if ((ops[i] & 0x7FFFFFFFu) >= K) return fail("op %u names topic %u >= K", i, ops[i] & 0x7FFFFFFFu);
  {  // the table keeps a free slot: at most cap - 1 keys at any time
    std::vector<u32> cnt(K, 0);
    u32 keys = 0;
    for (u32 i = 0; i < n_ops; ++i) {
      const u32 k = ops[i] & 0x7FFFFFFFu;
      if (ops[i] >> 31) {
        if (cnt[k] > 0 && --cnt[k] == 0) --keys;
      } else {
        if (cnt[k] == 0xFFFFu) return fail("op %u: count of topic %u past 65535", i, k);
        if (cnt[k]++ == 0 && ++keys >= cap) return fail("op %u: %u keys leave no free slot of %u", i, keys, cap);
      }
    }
  }
  Scratch sc;
  u32 *dops, *dout;
  CK(sc.upload(&dops, ops, n_ops));
  CK(sc.alloc(&dout, cap));
  k_stage_map<<<1, 64, (u64)doc_counts_bytes(K) + (u64)cap * 2>>>(dops, n_ops, cap, K, dout);
  HCK(hipGetLastError());
  HCK(hipMemcpy(table_out, dout, (u64)cap * 4, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int gpulda_stage_alias(const float* rows, uint32_t n_rows, uint32_t m, const uint64_t* divisor, float alpha,
                                  float* rows_out, float* sums_out, float* prob_out, uint32_t* alias_out) {
  Scratch sc;
  f32 *dr, *ds;
  unsigned long long* dd = nullptr;
  AliasSlot* da;
  const u64 total = (u64)n_rows * m;
  CK(sc.upload(&dr, rows, total));
  CK(sc.alloc(&ds, n_rows));
  CK(sc.alloc(&da, total));
  if (divisor) CK(sc.upload(&dd, reinterpret_cast<const unsigned long long*>(divisor), m));
  CK(launch_alias_rows(dr, n_rows, m, dd, ds, alpha, da, 0));
  HCK(hipDeviceSynchronize());
  if (rows_out) HCK(hipMemcpy(rows_out, dr, total * 4, hipMemcpyDeviceToHost));
  if (sums_out) HCK(hipMemcpy(sums_out, ds, (u64)n_rows * 4, hipMemcpyDeviceToHost));
  return split_alias(da, total, prob_out, alias_out);
}

extern "C" int gpulda_stage_alias_draw(const float* prob, const uint32_t* alias, uint32_t m, const uint32_t* r_slot,
                                       const uint32_t* r_keep, uint32_t count, uint32_t* out) {
  Scratch sc;
  AliasSlot* da;
  u32 *ds, *dk, *dout;
  CK(pack_alias(sc, prob, alias, m, &da));
  CK(sc.upload(&ds, r_slot, count));
  CK(sc.upload(&dk, r_keep, count));
  CK(sc.alloc(&dout, count));
  if (count) {
    k_stage_alias_draw<<<(count + 255) / 256, 256>>>(da, m, ds, dk, count, dout);
    HCK(hipGetLastError());
    HCK(hipMemcpy(out, dout, (u64)count * 4, hipMemcpyDeviceToHost));
  }
  return 0;
}

extern "C" int gpulda_stage_pick(const uint32_t* table, uint32_t cap, uint32_t K, const float* phi_row,
                                 float sigma_a, const float* u_decide, const float* u_target, uint32_t count,
                                 int use_regs, uint32_t* slot_out) {
  if (cap == 0 || cap % kMapLine) return fail("map capacity must be a positive multiple of 64");
  if (K == 0 || K > kMaxSweepTopics) return fail("K = %u out of range (1..%u)", K, kMaxSweepTopics);
  if ((u64)doc_counts_bytes(K) + (u64)cap * 2 > (u64)max_lds_per_block()) return fail("map capacity exceeds LDS");
  {  // occupied slots name distinct topics below K (the device state holds one count per topic)
    std::vector<u8> seen(K, 0);
    for (u32 s = 0; s < cap; ++s) {
      if (map_count(table[s]) == 0) continue;
      if (map_key(table[s]) >= K) return fail("slot %u names topic %u >= K", s, map_key(table[s]));
      if (seen[map_key(table[s])]++) return fail("slot %u repeats topic %u", s, map_key(table[s]));
    }
  }
  if (use_regs && cap / kMapLine > 8) return fail("register variants hold at most 8 lines");
  Scratch sc;
  u32 *dt, *dout;
  f32 *drow, *dux, *duw;
  CK(sc.upload(&dt, table, cap));
  CK(sc.upload(&drow, phi_row, K));
  CK(sc.upload(&dux, u_decide, count));
  CK(sc.upload(&duw, u_target, count));
  CK(sc.alloc(&dout, count));
  if (count) {
    k_stage_pick<<<count, 64, (u64)doc_counts_bytes(K) + (u64)cap * 2>>>(dt, cap, K, drow, sigma_a, dux, duw, count,
                                                                          use_regs ? 1u : 0u, dout);
    HCK(hipGetLastError());
    HCK(hipMemcpy(slot_out, dout, (u64)count * 4, hipMemcpyDeviceToHost));
  }
  return 0;
}

// Offsets of the documents of a slice as the sweeps read them (compute_d_idx, topics.cu:5-32):
// out[d] = first token of document d, out[n_docs] = total tokens.
extern "C" int gpulda_stage_doc_offsets(const uint32_t* d_len, uint32_t n_docs, uint64_t* out) {
  u64 N = 0;
  for (u32 i = 0; i < n_docs; ++i) N += d_len[i];
  std::vector<u32> w(std::max<u64>(N, 1), 0u);
  Segment* sg = new Segment();
  struct Guard {
    Segment* s;
    ~Guard() { free_segment(s); }
  } guard{sg};
  CK(build_segment(sg, w.data(), d_len, n_docs, N, 1, 1));
  HCK(hipMemcpy(out, sg->doc_off, ((u64)n_docs + 1) * sizeof(u64), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int gpulda_stage_poisson(const uint32_t* n, uint32_t count, float beta, uint32_t c1, uint32_t c2,
                                    uint32_t tag, uint64_t seed, uint32_t* out) {
  Scratch sc;
  AliasSlot* tables;
  u32 *dn, *dout;
  CK(stage_tables(sc, beta, &tables));
  CK(sc.upload(&dn, n, count));
  CK(sc.alloc(&dout, count));
  if (count) {
    k_stage_poisson<<<(count + 255) / 256, 256>>>(dn, count, beta, tables, c1, c2, tag, (u32)seed, (u32)(seed >> 32),
                                                 dout);
    HCK(hipGetLastError());
    HCK(hipMemcpy(out, dout, (u64)count * 4, hipMemcpyDeviceToHost));
  }
  return 0;
}

extern "C" int gpulda_stage_phi_draw(const uint32_t* n, uint32_t V, uint32_t K, float beta, uint32_t iter,
                                     uint64_t seed, float* phi_tilde, uint64_t* topic_sum) {
  Scratch sc;
  AliasSlot* tables;
  u32* dn;
  f32* dphi;
  unsigned long long* dts;
  const u64 VK = (u64)V * K;
  CK(stage_tables(sc, beta, &tables));
  CK(sc.upload(&dn, n, VK));
  CK(sc.alloc(&dphi, VK));
  CK(sc.alloc(&dts, K));
  CK(launch_phi_draw(dn, dphi, dts, V, K, beta, iter, nullptr, seed, tables, 0));
  HCK(hipDeviceSynchronize());
  HCK(hipMemcpy(phi_tilde, dphi, VK * 4, hipMemcpyDeviceToHost));
  HCK(hipMemcpy(topic_sum, dts, (u64)K * 8, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int gpulda_stage_sweep(uint32_t* z, const uint32_t* w, const uint64_t* doc_off, uint32_t n_docs,
                                  uint32_t doc_base, uint32_t iter, uint64_t seed, uint32_t K, uint32_t V,
                                  const float* phi, const float* sigma_a, const float* prob, const uint32_t* alias,
                                  uint64_t* visited) {
  if (K == 0 || K > kMaxSweepTopics) return fail("bad K");
  const u64 N = doc_off[n_docs];
  std::vector<u32> lens(n_docs);
  for (u32 d = 0; d < n_docs; ++d) {
    u64 L = doc_off[d + 1] - doc_off[d];
    if (L > kMaxDocLen) return fail("document %u longer than %u tokens", d, kMaxDocLen);
    lens[d] = (u32)L;
  }
  for (u64 i = 0; i < N; ++i)
    if (z[i] >= K) return fail("token %llu has z >= K", (unsigned long long)i);
  const u64 VK = (u64)V * K;
  Segment* sg = new Segment();
  int rc = build_segment(sg, w, lens.data(), n_docs, N, V, K);
  if (rc) {
    free_segment(sg);
    return rc;
  }
  struct Guard {
    Segment* s;
    ~Guard() { free_segment(s); }
  } guard{sg};
  Scratch sc;
  u32* dz32;
  f32 *dphi, *dsig;
  AliasSlot* dal;
  unsigned long long* dvis;
  u32* bad;
  CK(sc.upload(&dz32, z, N));
  CK(sc.upload(&dphi, phi, VK));
  CK(sc.upload(&dsig, sigma_a, V));
  CK(pack_alias(sc, prob, alias, VK, &dal));
  CK(sc.alloc(&dvis, 1));
  CK(sc.alloc(&bad, 1));
  CK(fill_u64(dvis, 1, 0, 0));
  CK(fill_u32(bad, 1, 0, 0));
  if (N) {
    k_z_narrow<<<grid_for(N, 256), 256>>>(dz32, sg->z, N, K, bad);
    HCK(hipGetLastError());
  }
  SweepParams P;
  P.w = sg->w;
  P.z = sg->z;
  P.doc_off = sg->doc_off;
  P.doc_order = nullptr;
  P.n_docs = n_docs;
  P.doc_base = doc_base;
  P.iter = iter;
  P.iter_ptr = nullptr;
  P.k0 = (u32)seed;
  P.k1 = (u32)(seed >> 32);
  P.K = K;
  P.phi = dphi;
  P.sigma_a = dsig;
  P.alias = dal;
  P.visited = dvis;
  CK(launch_sweep(P, sg, 0));
  if (N) {
    k_z_widen<<<grid_for(N, 256), 256>>>(sg->z, dz32, N);
    HCK(hipGetLastError());
  }
  HCK(hipDeviceSynchronize());
  if (N) HCK(hipMemcpy(z, dz32, N * 4, hipMemcpyDeviceToHost));
  if (visited) HCK(hipMemcpy(visited, dvis, 8, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int gpulda_stage_count(const uint32_t* w, const uint32_t* z, uint64_t n_tokens, uint32_t V, uint32_t K,
                                  uint32_t* n_out) {
  if (K == 0 || K > kMaxTopics) return fail("bad K");
  for (u64 i = 0; i < n_tokens; ++i)
    if (z[i] >= K) return fail("token %llu has z >= K", (unsigned long long)i);
  if (n_tokens >= (1ull << 32)) return fail("too many tokens");
  // the reduction does not look at documents: cut the tokens into maximal-length documents
  std::vector<u32> lens;
  for (u64 left = n_tokens; left > 0;) {
    u32 l = (u32)std::min<u64>(left, kMaxDocLen);
    lens.push_back(l);
    left -= l;
  }
  Segment* sg = new Segment();
  int rc = build_segment(sg, w, lens.data(), (u32)lens.size(), n_tokens, V, K);
  if (rc) {
    free_segment(sg);
    return rc;
  }
  Scratch sc;
  u32 *dz32, *dn, *bad;
  const u64 VK = (u64)V * K;
  int r2 = sc.upload(&dz32, z, n_tokens);
  if (!r2) r2 = sc.alloc(&dn, VK);
  if (!r2) r2 = sc.alloc(&bad, 1);
  if (!r2) r2 = fill_u32(dn, VK, 0, 0);
  if (!r2) r2 = fill_u32(bad, 1, 0, 0);
  if (!r2 && n_tokens) {
    k_z_narrow<<<grid_for(n_tokens, 256), 256>>>(dz32, sg->z, n_tokens, K, bad);
    if (hipGetLastError() != hipSuccess) r2 = fail("launch failed");
  }
  if (!r2) r2 = launch_count(sg, dn, K, 0);
  if (!r2 && hipDeviceSynchronize() != hipSuccess) r2 = fail("count failed");
  if (!r2 && hipMemcpy(n_out, dn, VK * 4, hipMemcpyDeviceToHost) != hipSuccess) r2 = fail("copy failed");
  free_segment(sg);
  return r2;
}

extern "C" int gpulda_stage_foldin(const uint32_t* w, const uint64_t* doc_off, uint32_t n_req, const uint32_t* req_ids,
                                   uint32_t sweeps, uint64_t seed, uint32_t K, uint32_t V, const float* phi,
                                   const float* sigma_a, const float* prob, const uint32_t* alias, uint32_t* z_out,
                                   uint64_t* visited) {
  if (K == 0 || K > kMaxSweepTopics) return fail("bad K");
  const u64 N = doc_off[n_req];
  std::vector<u32> lens(n_req);
  for (u32 r = 0; r < n_req; ++r) {
    u64 L = doc_off[r + 1] - doc_off[r];
    if (L > kMaxDocLen) return fail("request %u longer than %u tokens", r, kMaxDocLen);
    lens[r] = (u32)L;
  }
  for (u64 i = 0; i < N; ++i)
    if (w[i] >= V) return fail("word id >= V");
  const u64 VK = (u64)V * K;
  Scratch sc;
  u32 *dw, *dlen, *drid;
  u16* dz;
  u64* doff;
  f32 *dphi, *dsig;
  AliasSlot* dal;
  unsigned long long* dvis;
  CK(sc.upload(&dw, w, N));
  CK(sc.alloc(&dz, N));
  CK(sc.upload(&doff, reinterpret_cast<const u64*>(doc_off), n_req));
  CK(sc.upload(&dlen, lens.data(), n_req));
  CK(sc.upload(&drid, req_ids, n_req));
  CK(sc.upload(&dphi, phi, VK));
  CK(sc.upload(&dsig, sigma_a, V));
  CK(pack_alias(sc, prob, alias, VK, &dal));
  CK(sc.alloc(&dvis, 1));
  CK(fill_u64(dvis, 1, 0, 0));
  SweepParams P;
  P.w = dw;
  P.z = dz;
  P.doc_off = nullptr;
  P.doc_order = nullptr;
  P.n_docs = n_req;
  P.doc_base = 0;
  P.iter = 0;
  P.iter_ptr = nullptr;
  P.k0 = (u32)seed;
  P.k1 = (u32)(seed >> 32);
  P.K = K;
  P.phi = dphi;
  P.sigma_a = dsig;
  P.alias = dal;
  P.visited = dvis;
  FoldParams F;
  F.slots = nullptr;
  F.n_slots = n_req;
  F.span_off = doff;
  F.span_len = dlen;
  F.req_id = drid;
  F.sweeps = sweeps;
  F.z_out = nullptr;
  F.done = nullptr;
  F.done_value = 0;
  u32 maxL = 0;
  for (u32 L : lens) maxL = std::max(maxL, L);
  const u64 lds = doc_state_bytes(maxL, K);
  if (lds > (u64)max_lds_per_block()) return fail("doc-topic map exceeds LDS");
  if (n_req) {
    k_foldin<8><<<n_req, 64, lds>>>(P, F);
    HCK(hipGetLastError());
  }
  std::vector<u16> hz(N);
  HCK(hipDeviceSynchronize());
  if (N) HCK(hipMemcpy(hz.data(), dz, N * 2, hipMemcpyDeviceToHost));
  for (u64 i = 0; i < N; ++i) z_out[i] = hz[i];
  if (visited) HCK(hipMemcpy(visited, dvis, 8, hipMemcpyDeviceToHost));
  return 0;
}

// ---- allocation-free fold-in ------------------------------------------------------------------
namespace {
__global__ void k_span_lengths(const u64* __restrict__ off, u32 n, u32* __restrict__ len) {
  const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) len[i] = (u32)(off[i + 1] - off[i]);
}

// Carves 256-byte aligned pieces off the working-memory array.
struct Carver {
  unsigned char* base;
  u64 used = 0;
  template <class T>
  T* take(u64 count) {
    T* p = reinterpret_cast<T*>(base + used);
    used += (count * sizeof(T) + 255) & ~(u64)255;
    return p;
  }
};

u64 noalloc_bytes(u64 n_tokens, u64 n_req) {
  auto r = [](u64 b) { return (b + 255) & ~(u64)255; };
  return r(n_tokens * 4) + r(n_tokens * 2) + r((n_req + 1) * 8) + r(n_req * 4) + r(n_req * 4) + r(8);
}
}  // namespace

// The fold-in of gpulda_stage_foldin (same kernel, same results) against the library's current
// model -- phi, sigma_a and the alias tables of the last Phi step, and the library's seed -- with
// no device allocation: words, topics, offsets, lengths, request ids and the visit counter are
// staged through the static working-memory array.  Host arrays in and out; synchronous.
extern "C" int gpulda_foldin_noalloc(const uint32_t* w, const uint64_t* doc_off, uint32_t n_req,
                                     const uint32_t* req_ids, uint32_t sweeps, uint32_t* z_out, uint64_t* visited) {
  CK(require_ready());
  if (S.iter == 0) return fail("foldin_noalloc needs a trained model (run at least one Phi step)");
  if (!doc_off || (n_req && (!req_ids || !z_out))) return fail("foldin_noalloc: missing array");
  const u64 N = doc_off[n_req];
  u32 maxL = 0;
  for (u32 r = 0; r < n_req; ++r) {
    if (doc_off[r + 1] < doc_off[r]) return fail("request offsets decrease at %u", r);
    const u64 L = doc_off[r + 1] - doc_off[r];
    if (L > kMaxDocLen) return fail("request %u longer than %u tokens", r, kMaxDocLen);
    maxL = std::max(maxL, (u32)L);
  }
  if (doc_off[0] != 0) return fail("request offsets must start at 0");
  if (N && !w) return fail("foldin_noalloc: missing words");
  for (u64 i = 0; i < N; ++i)
    if (w[i] >= S.V) return fail("word id %u at token %llu is >= V = %u", w[i], (unsigned long long)i, S.V);
  const u64 need = noalloc_bytes(N, n_req);
  if (need > kNoAllocScratchBytes)
    return fail("foldin_noalloc: %llu tokens in %u requests need %llu bytes of working memory (limit %llu: 6 bytes "
                "per token, 16 per request)",
                (unsigned long long)N, n_req, (unsigned long long)need, (unsigned long long)kNoAllocScratchBytes);
  const u64 lds = doc_state_bytes(maxL, S.K);
  if (lds > (u64)max_lds_per_block()) return fail("doc-topic map exceeds LDS");
  unsigned char* base = nullptr;
  CK(scratch_on_current_device(&base));
  Carver cv{base};
  u32* dw = cv.take<u32>(N);
  u16* dz = cv.take<u16>(N);
  u64* doff = cv.take<u64>((u64)n_req + 1);
  u32* dlen = cv.take<u32>(n_req);
  u32* drid = cv.take<u32>(n_req);
  unsigned long long* dvis = cv.take<unsigned long long>(1);
  hipStream_t st = S.stream;
  if (N) HCK(hipMemcpyAsync(dw, w, N * sizeof(u32), hipMemcpyHostToDevice, st));
  HCK(hipMemcpyAsync(doff, doc_off, ((u64)n_req + 1) * sizeof(u64), hipMemcpyHostToDevice, st));
  if (n_req) {
    HCK(hipMemcpyAsync(drid, req_ids, (u64)n_req * sizeof(u32), hipMemcpyHostToDevice, st));
    k_span_lengths<<<(n_req + 255) / 256, 256, 0, st>>>(doff, n_req, dlen);
    HCK(hipGetLastError());
  }
  CK(fill_u64(dvis, 1, 0, st));  // the array keeps the previous call's contents
  SweepParams P = serve_params();
  P.w = dw;
  P.z = dz;
  P.n_docs = n_req;
  P.visited = dvis;
  FoldParams F;
  F.slots = nullptr;
  F.n_slots = n_req;
  F.span_off = doff;
  F.span_len = dlen;
  F.req_id = drid;
  F.sweeps = sweeps;
  F.z_out = nullptr;
  F.done = nullptr;
  F.done_value = 0;
  if (n_req) {
    k_foldin<8><<<n_req, 64, lds, st>>>(P, F);
    HCK(hipGetLastError());
  }
  if (N) {
    // the words are no longer needed: their space takes the topics widened to u32
    k_z_widen<<<grid_for(N, 256), 256, 0, st>>>(dz, dw, N);
    HCK(hipGetLastError());
    HCK(hipMemcpyAsync(z_out, dw, N * sizeof(u32), hipMemcpyDeviceToHost, st));
  }
  unsigned long long hv = 0;
  HCK(hipMemcpyAsync(&hv, dvis, sizeof(hv), hipMemcpyDeviceToHost, st));
  HCK(hipStreamSynchronize(st));
  if (visited) *visited = hv;
  return 0;
}
