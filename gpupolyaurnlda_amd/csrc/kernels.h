// kernels.h -- device kernel declarations and their parameter blocks.
#pragma once

#include "lda_core.h"

namespace lda {

// One word-sorted run of tokens for the n_wk reduction.
struct CountTask {
  u64 p0;    // first position in word-sorted order (the topics of the run are zs[p0 .. p0 + len))
  u32 len;   // number of tokens (<= kCountTask)
  u32 word;  // word id of the run
};
constexpr u32 kCountTask = 4096;
// Tokens per tile of the word-order passes (one workgroup: the tile's topics staged in 64 KiB of
// LDS), and the word groups of the first pass.
constexpr u32 kCountTile = 32768;
constexpr u32 kCountGroups = 256;

struct SweepParams {
  const u32* w;          // word ids, doc-major
  u16* z;                // topic indicators (read, then overwritten)
  const u64* doc_off;    // n_docs + 1 offsets into w / z
  const u32* doc_order;  // optional launch order (block -> doc); null = identity
  u32 n_docs;
  u32 doc_base;          // global id of document 0 (Philox counter word 1)
  u32 iter;              // iteration (Philox counter word 2) ...
  const u32* iter_ptr;   // ... or read from device memory when not null (replayable graphs)
  u32 k0, k1;            // Philox key = seed
  u32 K;
  const f32* phi;        // V x K word-major
  const f32* sigma_a;    // V
  const AliasSlot* alias;  // V x K word-major
  unsigned long long* visited;  // optional: += map entries visited by the sparse sums
  // Rows of the phi~ Poisson counts as u16 (half the bytes of the f32 rows), from which
  // phi = RN(count / S_k) is recovered exactly (phi_from_count); read only by
  // k_sweep_persistent_counts, for the words whose counts all fit 16 bits.
  const u16* phi_counts = nullptr;      // V x K, word-major (null: f32 rows only)
  const u8* counts_fit = nullptr;       // V: 1 when the row's counts all fit 16 bits
  const float2* topic_recip = nullptr;  // K + 1: (RN(1 / S_k), S_k); (0, 1) for S_k = 0 and entry K
};

// Fold-in batch: request slots of a token arena.
struct FoldParams {
  const u32* slots;     // batch -> slot (null: block b is slot b)
  u32 n_slots;
  const u64* span_off;  // slot -> first token in the arena
  const u32* span_len;  // slot -> tokens
  const u32* req_id;    // slot -> request id (Philox counter word 1)
  u32 sweeps;
  u16* z_out;           // optional copy of the final topics (arena offsets) the host reads
  u32* done;            // optional host-visible completion flags (slot)
  u32 done_value;
};

template <int LMAX>
__global__ void k_foldin(SweepParams P, FoldParams F);

// Request ring of the serving worker pool (device memory).
struct ServeRing {
  u32* ring;            // entry i % ring_size: slot + 1 of the i-th request once published, else 0
  u32 ring_size;        // = number of slots
  u32* tail;            // requests announced (an announced entry may still be being filled)
  u32* head;            // requests claimed
  u32* shutdown;        // nonzero asks the pool to leave
  u64 lifetime_ticks;   // wall-clock ticks after which a wave stops claiming or waiting
  u32* rec;             // per slot, kServeRecWords: the request record (see below)
  u64 arena;            // tokens in the arena (P.w / P.z)
  u32 V;                // word ids must lie below V
  u32* done_rec;        // host-visible, per slot kServeRecWords: what the serving wave read and wrote
};

// Request record, one 128-byte line per slot:
//   [0] ring index + 1 it was published at   [1] L   [2] request id   [3] offset low
//   [4] offset high   [5] sum of w_i * (2i + 1)
// A worker uses a record only once [0] names the entry it claimed and the words in the arena
// sum to [5] (both re-read for a bounded time); otherwise it reports the request as failed.
constexpr u32 kServeRecWords = 32;
constexpr u32 kServeMetaWords = 6;  // the part of a record the host sends ([0] holds the slot)
constexpr u32 kServeFailed = 2;     // completion flag of a request whose data never validated

// One flush.  The host has copied (by DMA, on the feeder's stream) each request's words into the
// arena and its record into meta; the feeder writes the records into the per-slot lines and
// publishes the ring entries.
struct ServeFeed {
  const u32* meta;      // kServeMetaWords per request of this flush, in ring order
  u32 base;             // ring index of the first record
  u32 count;
};
__global__ void k_serve_feed(ServeFeed Fd, ServeRing R);
__global__ void k_serve_worker(SweepParams P, FoldParams F, ServeRing R);

__global__ void k_fill_u32(u32* p, u64 n, u32 v);
__global__ void k_fill_u64(unsigned long long* p, u64 n, unsigned long long v);
__global__ void k_z_narrow(const u32* in, u16* out, u64 n, u32 K, u32* bad);
__global__ void k_z_widen(const u16* in, u32* out, u64 n);
__global__ void k_z_init(u16* z, const u64* doc_off, u32 n_docs, u32 doc_base, u32 K, u32 k0, u32 k1);
// counts (optional): the draws as u16 rows (clamped), wide[v] set where a draw of row v passes 16 bits.
__global__ void k_phi_draw(const u32* n, f32* phi, unsigned long long* topic_sum, u32 V, u32 K, f32 beta, u32 iter,
                           const u32* iter_ptr, u32 k0, u32 k1, const AliasSlot* tables, u32 rows_per_block,
                           u16* counts, u8* wide);
__global__ void k_bump_u32(u32* p);
__global__ void k_init_counts(const u32* C, u32* n, u32 V, u32 K, f32 beta, u32 k0, u32 k1, const AliasSlot* tables,
                              u32 rows_per_block);
template <int PER>
__global__ void k_alias_rows(f32* rows, u32 n_rows, u32 m, const unsigned long long* divisor, f32* sigma, f32 alpha,
                             AliasSlot* out, u32 rows_per_block_waves);
template <int LMAX>
__global__ void k_sweep_lines(SweepParams P);
template <int LMAX>
__global__ void k_sweep_persistent(SweepParams P, u32* head, u32 n_items);
__global__ void k_sweep_persistent_long(SweepParams P, u32* head, u32 n_items);
// Persistent sweep reading the u16 count rows where they fit: blocks of kCountsBlockWaves
// wavefronts share one LDS copy of topic_recip; wavefront w's document state lies at
// topic_recip_lds_bytes(K) + w * wave_lds bytes.
constexpr u32 kCountsBlockWaves = 16;
LDA_HD u64 topic_recip_lds_bytes(u32 K) { return ((u64)(K + 1) * 8 + 15) / 16 * 16; }
__global__ void k_sweep_persistent_counts(SweepParams P, u32* head, u32 n_items, u32 wave_lds);
__global__ void k_topic_recip(const unsigned long long* topic_sum, float2* out, u32 K);
// Number of cells of the rows marked as fitting whose recovered phi differs from the f32 row.
__global__ void k_check_counts(const f32* phi, const u16* counts, const u8* fit, const float2* recip, u32 V, u32 K,
                               unsigned long long* bad);
// Block-per-document sweep (4 wavefronts per document); head == null: block b takes item b.
__global__ void k_sweep_block(SweepParams P, u32* head, u32 n_items);
// Launch buckets of the per-bucket schedule, by map lines (0 = any size).  Buckets 1-8 run the
// kernel with that many register-resident lines; 16, 32 and 0 run the line-loop kernel
// (k_sweep_lines<0>) with the LDS of their largest map.
constexpr int kSweepBuckets[] = {1, 2, 4, 8, 16, 32};
// Kernel family member for a map of `lines` 64-slot lines.
inline int sweep_bucket_lmax(u32 lines) {
  for (int b : kSweepBuckets)
    if (lines <= (u32)b) return b;
  return 0;
}
// Output runs of the kCountTile-entry tiles of one reorder pass (built once per corpus slice).
struct TileRuns {
  u16* rank = nullptr;     // per input entry: its rank inside its tile (output order)
  u32* bits = nullptr;     // per tile, kCountTile / 32 words: bit r set when rank r starts a run
  u32* run_off = nullptr;  // per tile (n_tiles + 1 entries): index of its first run
  uint2* runs = nullptr;   // per run: {first rank, place in the output of that rank}
};
__global__ void k_tile_reorder(const u16* in, TileRuns R, u16* out, u64 n);
__global__ void k_count(const CountTask* tasks, u32 n_tasks, const u16* zs, u32* n, u32 K, u32 waves_per_block);
__global__ void k_count_direct(const CountTask* tasks, u32 n_tasks, const u16* zs, u32* n, u32 K);

__global__ void k_stage_philox(const u32* ctr, const u32* key, u32* out, u32 n);
__global__ void k_stage_scan(const f32* in, f32* out, u32 n_groups);
__global__ void k_stage_map(const u32* ops, u32 n_ops, u32 cap, u32 K, u32* out);
__global__ void k_stage_poisson(const u32* nval, u32 count, f32 beta, const AliasSlot* tables, u32 c1, u32 c2,
                                u32 tag, u32 k0, u32 k1, u32* out);
__global__ void k_stage_pick(const u32* table, u32 cap, u32 K, const f32* phi_row_k, f32 sa, const f32* ux,
                             const f32* uw, u32 count, u32 use_regs, u32* out);
__global__ void k_stage_alias_draw(const AliasSlot* row, u32 m, const u32* r_slot, const u32* r_keep, u32 count,
                                   u32* out);

// LDS bytes per alias row of length m (one wavefront).
LDA_HD u64 alias_lds_per_wave(u32 m) { return (u64)((m + 3) & ~3u) * 14 + 68 * 4; }
// Block-per-row alias kernel (kAliasBlockThreads threads per row, rows of up to 256 * PER slots):
// dynamic LDS per row (u64 prefixes, padded f32 row values, u16 slot indices).
constexpr u32 kAliasBlockThreads = 256;
LDA_HD u64 alias_block_lds(u32 m) { return (u64)((m + 3) & ~3u) * 14 + (kAliasBlockThreads + 4) * 4; }
// fit_out (optional) receives whether the sweep is to read each row's u16 counts (written by
// k_phi_draw): none passed 16 bits (wide, cleared here for the next step) and keep_f32 (optional)
// does not mark the row.  recip (optional, with divisor) holds (RN(1 / d), d) per slot for
// d = (f32)divisor (k_topic_recip); given only for rows of integer counts (the Poisson draws).
template <int PER>
__global__ void k_alias_block(f32* rows, u32 n_rows, u32 m, const unsigned long long* divisor, f32* sigma, f32 alpha,
                              AliasSlot* out, u8* wide, u8* fit_out, const u8* keep_f32, const float2* recip);

}  // namespace lda
