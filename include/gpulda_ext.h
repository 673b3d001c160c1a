/*
 * gpulda_ext.h -- extension entry points of libGPUPolyaUrnLDA.so.
 *
 * The reference's Rust host never calls these; they exist for callers that need more than the
 * five-call ABI of gpulda_abi.h: a seed, state read-back, a device-resident corpus (no per-call
 * PCIe traffic), stage-level entry points for conformance tests, and the fold-in server.
 * All return 0 on success and a nonzero code on failure; gpulda_last_error() describes the last
 * failure.  None of them terminates the process.
 *
 * Semantics shared with the ABI path (each fixes a reference defect; reference paths are
 * relative to the reference repository root):
 *   - Randomness is a counter-based Philox4x32-10 keyed by the seed (default 0) with counters
 *     (token-in-document, global document id, iteration, purpose).  The reference reuses one
 *     cuRAND state whose per-buffer copy is never advanced (src/cu/train.cu:168) and ignores
 *     --seed (src/args.rs:36-37).
 *   - Global document ids count documents in the order sample_z_async sees them within an
 *     iteration (train.rs walks the corpus in file order); iteration t is the t-th sample_phi.
 *   - The sparse term uses phi[w][k] (the reference reads phi[k] only, src/cu/topics.cuh:174).
 *   - Alias tables store the probability of keeping a slot (scaled convention; the reference
 *     mixes conventions, src/cu/spalias.cu:135 vs src/cu/topics.cuh:69-76).
 *   - All reductions are exact (integer) or use one fixed floating-point order, so results are
 *     bit-identical run to run and independent of launch shape.
 *   - Poisson draws are integers for every lambda (the reference's n >= 100 branch returns a
 *     non-integer Gaussian, src/cu/polyaurn.cu:30-35).
 *   - A token whose word has an all-zero phi row keeps its topic (the reference would read 0/0).
 *
 * Limits: K <= 4096 (LDS-resident doc-topic maps and alias split), documents <= 65535 tokens.
 *
 * Layouts: phi, n_wk and alias tables are word-major, element [v * K + k].
 */
#ifndef GPULDA_EXT_H
#define GPULDA_EXT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- configuration and diagnostics ------------------------------------------------------- */
int gpulda_set_seed(uint64_t seed);           /* before initialize(); default 0 */
uint64_t gpulda_get_seed(void);
const char* gpulda_last_error(void);
/* Device name, compute units, HBM bytes, max LDS bytes per workgroup of the current device. */
int gpulda_device_info(char* name, int name_len, int* cu_count, uint64_t* hbm_bytes, int* lds_per_block);
/* Device copies of buffer slices (w, d and what the sweep derives from them) across
 * sample_z_async calls.  The reference uploads w and d on every call (src/cu/train.cu:160-163).
 *   0  no cache: every call uploads its slice;
 *   1  (default) a slice is reused when its whole content matches -- every document length and
 *      word id, by a 64-bit hash computed on host threads at each call -- whatever its pointers;
 *   2  trusted: reused when the host pointers, sizes and a sampled fingerprint match, for callers
 *      whose w and d never change between calls (read-only mappings, as the reference's host).
 * Cached slices are held within a byte budget (default: half the free HBM at initialize), least
 * recently used first out; a slice a buffer has in flight is never evicted. */
int gpulda_set_slice_cache(int mode);
int gpulda_set_slice_cache_budget(uint64_t bytes);  /* 0 = the default budget */
/* Slices cached now, their device bytes, and slices built (uploads of w and d) so far. */
int gpulda_slice_cache_stats(uint32_t* segments, uint64_t* bytes, uint64_t* builds);
/* The content hash of cache mode 1 (host only; no device needed). */
uint64_t gpulda_slice_hash(const uint32_t* w, const uint32_t* d_len, uint32_t n_docs, uint64_t n_tokens);
/* sample_z_async splits a slice into up to 8 parts of whole documents and at least min_tokens
 * tokens each (default 4 Mi), and overlaps the copies of z for one part with the sweep of another.
 * Applies to slices first seen after the call.  Results do not depend on it. */
int gpulda_set_part_tokens(uint64_t min_tokens);
/* Count rows (default on): each Phi step also writes the rows of the phi~ Poisson counts as u16
 * for the words whose counts all fit 16 bits, and the persistent sweep reads those (half the bytes
 * of an f32 row) and recovers phi = count / S_k exactly.  Off: the f32 rows only.  The sampled
 * z does not depend on it.  Switching on takes effect at the next Phi step. */
int gpulda_set_count_rows(int on);
/* The n most frequent words (by the counts C given to initialize) keep f32 rows even when their
 * counts fit: their rows stay in L2, so the count rows would add conversion work and save no
 * misses.  Default: as many words as 1 MiB of f32 rows (256 at K = 1024).  From the next Phi step. */
int gpulda_set_f32_rows(uint32_t n_words);
/* Whether count rows are on, and (when on) the count rows of the last Phi step checked against the
 * f32 rows: cells whose recovered phi differs (bitwise), and the words whose rows the sweep reads
 * as counts. */
int gpulda_check_count_rows(int* on, uint64_t* mismatches, uint64_t* words_fit);

/* ---- model state (valid between initialize and cleanup) ---------------------------------- */
uint32_t gpulda_iteration(void);                         /* number of sample_phi calls so far */
int gpulda_copy_phi(float* out);                         /* V*K */
int gpulda_copy_alias(float* prob, uint32_t* alias);     /* V*K each */
int gpulda_copy_sigma_a(float* out);                     /* V */
int gpulda_copy_counts(uint32_t* out);                   /* n_wk, V*K */
int gpulda_set_counts(const uint32_t* n);                /* overwrite n_wk, V*K */
/* Number of sample_phi calls taken so far (the next one is iteration + 1).  With n_wk set to the
 * counts of a saved z, a resumed run continues the uninterrupted chain bit for bit. */
int gpulda_set_iteration(uint32_t iteration);
int gpulda_copy_topic_sums(uint64_t* out);               /* sum_v phi~_vk of the last Phi step, K */
int gpulda_copy_poisson_tables(float* prob, uint32_t* alias);  /* 100 x 200 each */
/* Doc-topic map entries read by the sparse sums since the last reset (sum over tokens). */
int gpulda_sweep_stats(uint64_t* visited);
int gpulda_reset_stats(void);

/* ---- device-resident corpus ------------------------------------------------------------- */
/* Uploads a whole corpus (word ids w, document lengths d_len) once; z == NULL draws the
 * initial topics on the device (Philox, uniform over K), otherwise z is uploaded. */
int gpulda_corpus_load(const uint32_t* w, const uint32_t* d_len, uint32_t n_docs, const uint32_t* z);
int gpulda_corpus_counts_from_z(void);                   /* n_wk := counts of the current z */
int gpulda_corpus_copy_z(uint32_t* z);
/* n_iters full iterations (Phi step, alias tables, z-sweep, n_wk reduction) on the device.
 * stage_ms (optional, 4 floats): mean device ms per iteration of
 * [Poisson draw, alias rows, z-sweep, n_wk reduction], timed with hipEvents on the library
 * stream. */
int gpulda_corpus_iterate(uint32_t n_iters, float* stage_ms);
/* Sweep schedule of the resident corpus (all produce bit-identical results):
 *   0  one launch per length bucket (wavefront per document; default);
 *   1  one persistent launch of cu_count * waves_per_cu wavefronts over a longest-first queue;
 *   2  as 1, plus a block-per-document persistent launch (4 wavefronts share a document) for
 *      documents of at least block_min_len tokens, on a second stream;
 *   3  two persistent launches: documents longer than block_min_len (LDS for the longest map)
 *      on a second stream, the rest with LDS sized for their own longest map (more wavefronts
 *      per CU). */
int gpulda_corpus_schedule(int schedule, uint32_t waves_per_cu, uint32_t block_min_len);
/* 1 = capture one iteration into a hipGraph and replay it (the iteration number lives in device
 * memory); stage_ms of gpulda_corpus_iterate then reports only the total, in stage_ms[2]. */
int gpulda_corpus_use_graph(int enabled);
/* Times one sweep under each candidate schedule on the resident corpus (z restored after each
 * trial) and applies the fastest.  With count rows on, each persistent schedule is also timed on
 * f32 rows, and count rows are switched off when that wins.  table: 5 floats per candidate
 * {schedule, waves_per_cu, block_min_len, count rows (0/1), ms}.  Needs at least one iteration
 * first. */
int gpulda_corpus_autotune(float* table, uint32_t max_candidates, uint32_t* n_candidates, int* chosen);
/* Collapsed joint log-likelihood log p(w, z | alpha, beta) of the current z and n_wk. */
int gpulda_corpus_loglik(double* out);
int gpulda_corpus_release(void);

/* Measurement aid: `reps` launches of a plain device-to-device copy kernel over `bytes` bytes
 * (reads `bytes`, writes `bytes`), used to scale hardware byte counters; mean ms per copy. */
int gpulda_calibrate_copy(uint64_t bytes, uint32_t reps, float* ms_per_copy);

/* ---- fold-in server (continuous batching against the resident model) ----------------------
 * Keeps phi, sigma_a and the alias tables of the current model in HBM.  Requests are admitted
 * while a slot and arena space (hbm_budget_bytes / 6 tokens) are free; flush publishes every
 * admitted request to a request ring that a resident pool of worker wavefronts (waves_per_cu per
 * compute unit, one wavefront per request) drains as requests arrive, so new requests start while
 * earlier ones are still running.  Each request runs `sweeps` passes, then its topics and
 * completion flag become visible in host memory (poll / result).  A request's randomness depends
 * on its id only: its result is the same alone or beside any other requests. */
int gpulda_serve_open(uint64_t hbm_budget_bytes, uint32_t sweeps, uint32_t waves_per_cu, uint32_t max_slots);
/* 0 = admitted (slot set), 1 = no room now (retry after results are collected). */
int gpulda_serve_submit(const uint32_t* w, uint32_t L, uint32_t request_id, uint32_t* slot);
int gpulda_serve_flush(void);
int gpulda_serve_poll(uint32_t* done_slots, uint32_t max, uint32_t* n_done);
int gpulda_serve_result(uint32_t slot, uint32_t* z, uint32_t* request_id);
int gpulda_serve_stats(uint32_t* pending, uint32_t* inflight, uint64_t* free_tokens);
/* A published request that gets no result within `seconds` (default 30) of its publication is
 * reported by gpulda_serve_poll and collected as failed by gpulda_serve_result; its slot and arena
 * span return once no worker kernel runs.  No reference counterpart (the reference has no server). */
int gpulda_serve_set_deadline(double seconds);
int gpulda_serve_close(void);

/* ---- stage entry points (host arrays in and out; the library stages device copies) ------- */
int gpulda_stage_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out, uint32_t n);
int gpulda_stage_scan64(const float* in, float* out, uint32_t n_groups);
/* ops[i] = topic (increment) or topic | 1u<<31 (decrement of a present topic), topics below
 * K <= 32767 (the home slots depend on K); the table keeps a free slot (at most cap - 1 topics at
 * any time, checked).  Returns the final table as entries key << 16 | count (0xFFFF0000 empty). */
int gpulda_stage_map(const uint32_t* ops, uint32_t n_ops, uint32_t cap, uint32_t K, uint32_t* table_out);
/* Row normalisation (by divisor[k] when divisor != NULL), row sums and alias tables. */
int gpulda_stage_alias(const float* rows, uint32_t n_rows, uint32_t m, const uint64_t* divisor, float alpha,
                       float* rows_out, float* sums_out, float* prob_out, uint32_t* alias_out);
/* draw_alias (src/cu/topics.cuh:63-83): slot floor(r_slot * m / 2^32), kept when
 * u(r_keep) < prob[slot], else alias[slot]. */
int gpulda_stage_alias_draw(const float* prob, const uint32_t* alias, uint32_t m, const uint32_t* r_slot,
                            const uint32_t* r_keep, uint32_t count, uint32_t* out);
/* Sparse/dense decision and sparse pick of the z-sweep on a given doc-topic map (cap slots of
 * entries key << 16 | count, count 0 for an empty slot, occupied slots naming distinct topics
 * below K <= 32767, checked) for a word whose phi row is phi_row (K values) and sigma_a > 0:
 * draw i is sparse when u_decide[i] * (sigma_a + sigma_b) > sigma_a and then picks the slot whose
 * interval holds u_target[i] * sigma_b (compute_product_cumsum + draw_wary_search,
 * src/cu/topics.cuh:163-192, 88-137).  slot_out[i] = that slot, or 0xFFFFFFFF for a dense draw.
 * use_regs selects the register variants (cap <= 512), else the line-loop form. */
int gpulda_stage_pick(const uint32_t* table, uint32_t cap, uint32_t K, const float* phi_row, float sigma_a,
                      const float* u_decide, const float* u_target, uint32_t count, int use_regs,
                      uint32_t* slot_out);
/* Document offsets as the sweeps read them (compute_d_idx, src/cu/topics.cu:5-32): out[d] = first
 * token of document d, out[n_docs] = total (n_docs + 1 values). */
int gpulda_stage_doc_offsets(const uint32_t* d_len, uint32_t n_docs, uint64_t* out);
/* Poisson draws with counters (i, c1, c2, tag) for i = 0..count-1, tables built for beta. */
int gpulda_stage_poisson(const uint32_t* n, uint32_t count, float beta, uint32_t c1, uint32_t c2, uint32_t tag,
                         uint64_t seed, uint32_t* out);
int gpulda_stage_phi_draw(const uint32_t* n, uint32_t V, uint32_t K, float beta, uint32_t iter, uint64_t seed,
                          float* phi_tilde, uint64_t* topic_sum);
int gpulda_stage_sweep(uint32_t* z, const uint32_t* w, const uint64_t* doc_off, uint32_t n_docs, uint32_t doc_base,
                       uint32_t iter, uint64_t seed, uint32_t K, uint32_t V, const float* phi, const float* sigma_a,
                       const float* prob, const uint32_t* alias, uint64_t* visited);
int gpulda_stage_count(const uint32_t* w, const uint32_t* z, uint64_t n_tokens, uint32_t V, uint32_t K,
                       uint32_t* n_out);
/* Fold-in of requests (documents of w delimited by doc_off, ids req_ids) against the given model. */
int gpulda_stage_foldin(const uint32_t* w, const uint64_t* doc_off, uint32_t n_req, const uint32_t* req_ids,
                        uint32_t sweeps, uint64_t seed, uint32_t K, uint32_t V, const float* phi,
                        const float* sigma_a, const float* prob, const uint32_t* alias, uint32_t* z_out,
                        uint64_t* visited);

/* ---- allocation-free fold-in ---------------------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc): the fold-in of gpulda_stage_foldin (same kernel, same
 * results) against the library's current model -- phi, sigma_a and the alias tables of the last
 * Phi step, and the seed set by gpulda_set_seed -- that calls no hipMalloc, hipFree or other
 * allocating routine.  Words, topics, offsets and counters are staged through one static device
 * array of GPULDA_NOALLOC_SCRATCH_BYTES on each GPU, held until the process ends; a call needing
 * more (6 bytes per token + 16 per request, each piece rounded to 256 bytes) is rejected with an
 * error.  Every call on a GPU shares that GPU's array: calls that overlap on one GPU race.
 * Requests are the documents of w delimited by doc_off (doc_off[0] = 0, n_req + 1 offsets);
 * z_out receives their topics, *visited the map entries the sparse sums read. */
#define GPULDA_NOALLOC_SCRATCH_BYTES (16u << 20)
int gpulda_foldin_noalloc(const uint32_t* w, const uint64_t* doc_off, uint32_t n_req, const uint32_t* req_ids,
                          uint32_t sweeps, uint32_t* z_out, uint64_t* visited);

#ifdef __cplusplus
}
#endif

#endif /* GPULDA_EXT_H */
