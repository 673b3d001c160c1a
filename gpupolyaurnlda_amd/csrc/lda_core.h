// lda_core.h -- arithmetic shared by the host side and the gfx950 kernels of libGPUPolyaUrnLDA.
//
// Everything here is compiled identically for the host (x86-64) and the device (gfx950) with
// -ffp-contract=off, uses only IEEE-754 +, -, *, / and integer operations (no transcendental
// library calls), and is therefore bit-reproducible on both sides.  The randomness of every
// sampling step is a pure function of a Philox4x32-10 counter (purpose, iteration, document,
// token) and the seed, so a result never depends on launch shape or scheduling.
//
// Reference counterparts:
//   philox4x32_10         replaces the cuRAND Philox state + skipahead of src/cu/random.cu:6-17,
//                         src/cu/topics.cu:41,48-50,69-70, src/cu/polyaurn.cu:65-66,91-92
//   draw_alias_slot       replaces draw_alias, src/cu/topics.cuh:63-83 (scaled-probability
//                         convention; the reference compares a scaled u' against unscaled probs)
//   poisson_draw          replaces draw_poisson, src/cu/polyaurn.cu:7-36 (tables for n < 100 as
//                         the reference; an exact integer PTRS draw instead of the non-integer
//                         Gaussian branch for n >= 100)
//   map_capacity          replaces the capacity rule of HashMap::allocate, src/cu/hashmap.cuh:98
#pragma once

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LDA_HD __host__ __device__ __forceinline__
#else
#define LDA_HD static inline
#endif

namespace lda {

typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int32_t i32;
typedef int64_t i64;
typedef float f32;
typedef double f64;

struct u32x4 {
  u32 x, y, z, w;
};

// ---------------------------------------------------------------------------------------------
// Philox4x32-10, Salmon, Moraes, Dror & Shaw, "Parallel Random Numbers: As Easy as 1, 2, 3",
// SC'11.  Multipliers and Weyl increments are the published constants.
// ---------------------------------------------------------------------------------------------
constexpr u32 kPhiloxM0 = 0xD2511F53u;
constexpr u32 kPhiloxM1 = 0xCD9E8D57u;
constexpr u32 kPhiloxW0 = 0x9E3779B9u;
constexpr u32 kPhiloxW1 = 0xBB67AE85u;

LDA_HD u32x4 philox_round(u32x4 c, u32 k0, u32 k1) {
  u64 p0 = (u64)kPhiloxM0 * (u64)c.x;
  u64 p1 = (u64)kPhiloxM1 * (u64)c.z;
  u32x4 r;
  r.x = (u32)(p1 >> 32) ^ c.y ^ k0;
  r.y = (u32)p1;
  r.z = (u32)(p0 >> 32) ^ c.w ^ k1;
  r.w = (u32)p0;
  return r;
}

LDA_HD u32x4 philox4x32_10(u32x4 c, u32 k0, u32 k1) {
#pragma unroll
  for (int i = 0; i < 9; ++i) {
    c = philox_round(c, k0, k1);
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
  return philox_round(c, k0, k1);
}

// Purpose tags occupy the top byte of counter word 3; the low 24 bits are a sub-stream index
// (the attempt number of a rejection sampler).
enum Tag : u32 {
  kTagInitZ = 1,      // initial topic indicators of a generated corpus
  kTagInitN = 2,      // initial word-topic counts drawn from corpus counts (initialize())
  kTagPhi = 3,        // Poisson Polya-urn draw of phi~ at iteration t
  kTagSweep = 4,      // z-sweep: one draw of four words per token
  kTagFoldInit = 5,   // fold-in: initial topic of a request token
  kTagFoldSweep = 6,  // fold-in: sweep s of a request
};

LDA_HD u32x4 philox_tagged(u32 c0, u32 c1, u32 c2, u32 tag, u32 sub, u32 k0, u32 k1) {
  u32x4 c;
  c.x = c0;
  c.y = c1;
  c.z = c2;
  c.w = (tag << 24) | (sub & 0x00FFFFFFu);
  return philox4x32_10(c, k0, k1);
}

// Uniform in [0, 1) on a 2^-24 grid (exact conversion; never returns 1).
LDA_HD f32 u24(u32 r) { return (f32)(r >> 8) * 0x1p-24f; }

// Uniform in [0, 1) on a 2^-53 grid from two words (exact in double).
LDA_HD f64 u53(u32 a, u32 b) {
  return ((f64)(a >> 5) * 67108864.0 + (f64)(b >> 6)) * 0x1p-53;
}

// Lemire multiply-shift: an index in [0, n) from 32 random bits.
LDA_HD u32 bounded(u32 r, u32 n) { return (u32)(((u64)r * (u64)n) >> 32); }

// ---------------------------------------------------------------------------------------------
// Doc-topic map (wave-cooperative Robin Hood table).  On the device a document's map is a 16-bit
// count per topic plus a table of the topics present (the sweep kernels' LDS state); as one table
// of entries key << 16 | count it is what the conformance stages exchange with the tests.  The
// table is history-independent: Robin Hood linear probing with ties broken by key, and deletion by
// backward shift, so its layout is a function of the current key set only.
// ---------------------------------------------------------------------------------------------
// In the entry form an empty slot holds count 0 at key 0xFFFF (kMapEmpty): a stored entry has a
// count of at least 1, so an empty slot's mass count * phi[w][key] is +0 whatever its key field.
constexpr u32 kMapEmpty = 0xFFFF0000u;
constexpr u32 kMapLine = 64;          // one wavefront probes 64 consecutive slots
constexpr u32 kMaxDocLen = 65535;     // counts are 16-bit
constexpr u32 kMaxTopics = 65536;     // keys are 16-bit
// The sweep's key table stores 2 * topic in 16-bit slots (the byte offset of the topic's count),
// and 2 * K for an empty slot.
constexpr u32 kMaxSweepTopics = 32767;

// Completion flag of a served request: a fingerprint of its words (sum of w_i * (2i + 1), any
// order), length and request id, never 0 (0 means not finished).
LDA_HD u32 serve_fingerprint(u32 word_sum, u32 L, u32 rid) {
  return (word_sum ^ (L * 0x9E3779B1u) ^ (rid * 0x85EBCA6Bu)) | 1u;
}

LDA_HD u32 map_key(u32 e) { return e >> 16; }
LDA_HD u32 map_count(u32 e) { return e & 0xFFFFu; }
LDA_HD u32 map_entry(u32 key, u32 count) { return (key << 16) | count; }
LDA_HD bool map_is_empty(u32 e) { return map_count(e) == 0u; }

// Home slot of a topic below K: key * hm with hm = floor((2^32 - 1) / K) stays below 2^32, and the
// multiply-shift maps it to [0, C), non-decreasing in the key.  A line of the table then holds a
// run of neighbouring topics, so the phi values its 64 slots gather lie in a few cache lines
// (a multiplicative hash spreads them over the whole phi row, one cache line per slot).
LDA_HD u32 map_home_mult(u32 K) { return 0xFFFFFFFFu / K; }
// floor(key * hm * cap / 2^32), formed as (key * cap) * hm: the same integer product, with key * cap
// a 24-bit multiply on the device.  Needs key * cap < 2^32, which K <= 32767 (cap <= 33,792) gives.
LDA_HD u32 map_home(u32 key, u32 cap, u32 hm) {
  return (u32)(((u64)((key & 0xFFFFFFu) * (cap & 0xFFFFFFu)) * (u64)hm) >> 32);
}

// Displacement of an entry with home h stored in slot s.
LDA_HD u32 map_disp(u32 s, u32 h, u32 cap) { return s >= h ? s - h : s + cap - h; }

// Slots for a document of length L over K topics: room for every distinct topic it can hold,
// load factor <= 32/33, at least one free slot, rounded up to whole 64-slot lines.  Dense tables
// keep the lines each token's sum reads few; Robin Hood order keeps their probes short.
LDA_HD u32 map_capacity(u32 L, u32 K) {
  u32 need = L < K ? L : K;
  u32 c = need + need / 32 + 1;
  return ((c + kMapLine - 1) / kMapLine) * kMapLine;
}

// LDS of one document's sweep state: K + 1 counts (u16; entry K stays 0), padded to 8 bytes, then
// the key table (u16 per slot) for the largest map of a document of up to maxL tokens.
LDA_HD u32 doc_counts_bytes(u32 K) { return ((K + 1) * 2 + 7) / 8 * 8; }
LDA_HD u64 doc_state_bytes(u32 maxL, u32 K) { return (u64)doc_counts_bytes(K) + (u64)map_capacity(maxL, K) * 2; }

// ---------------------------------------------------------------------------------------------
// Alias tables.  prob is the probability of keeping the slot (scaled convention, in [0, 1]).
// ---------------------------------------------------------------------------------------------
struct AliasSlot {
  f32 prob;
  u32 alias;
};

// phi = RN(x / d) for a phi~ count x < 2^16 and a topic divisor d = (f32)S_k > 0, from
// y = RN(1 / d): q0 = RN(x * y) is within an ulp of x / d, its remainder x - q0 * d is exact in
// one fma, and the correction q0 + r * y rounds to RN(x / d) (Markstein's theorem for a correctly
// rounded reciprocal).  Every operation scales exactly with powers of two here (no underflow or
// overflow for x < 2^16, 1 <= d < 2^64), so checking every 16-bit significand of x against every
// 24-bit significand of d covers all inputs; x = 0 gives +0, and so does y = 0 with d = 1 (the
// value used for S_k = 0, whose counts are all 0).
LDA_HD f32 phi_from_count_f(f32 xf, f32 y, f32 d) {  // xf: a count below 2^16, as f32
  const f32 q0 = xf * y;
  const f32 r = __builtin_fmaf(-q0, d, xf);
  return __builtin_fmaf(r, y, q0);
}
LDA_HD f32 phi_from_count(u32 x, f32 y, f32 d) { return phi_from_count_f((f32)x, y, d); }

LDA_HD u32 draw_alias_slot(const AliasSlot* row, u32 n, u32 r_slot, u32 r_keep) {
  u32 slot = bounded(r_slot, n);
  AliasSlot s = row[slot];
  return u24(r_keep) < s.prob ? slot : s.alias;
}

// ---------------------------------------------------------------------------------------------
// Deterministic transcendental helpers (double).  Written from the textbook formulas so that
// the host and the device evaluate exactly the same operations.
// ---------------------------------------------------------------------------------------------
LDA_HD u64 f64_bits(f64 x) {
  u64 b;
  memcpy(&b, &x, 8);
  return b;
}
LDA_HD f64 f64_from_bits(u64 b) {
  f64 x;
  memcpy(&x, &b, 8);
  return x;
}

// Natural log for finite x > 0: x = 2^e * m with m in (sqrt(1/2), sqrt(2)], then
// log m = 2 atanh(f), f = (m-1)/(m+1), |f| <= 0.1716, by the odd series to f^21
// (truncation < 2^-53 relative).  ln 2 split in a high and a low part (Cody & Waite).
LDA_HD f64 det_log(f64 x) {
  u64 b = f64_bits(x);
  int e = (int)((b >> 52) & 0x7FF);
  if (e == 0) {  // subnormal input: scale into the normal range
    x = x * 0x1p54;
    b = f64_bits(x);
    e = (int)((b >> 52) & 0x7FF) - 54;
  }
  e -= 1023;
  f64 m = f64_from_bits((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  f64 f = (m - 1.0) / (m + 1.0);
  f64 f2 = f * f;
  f64 p = 1.0 / 21.0;
  p = p * f2 + 1.0 / 19.0;
  p = p * f2 + 1.0 / 17.0;
  p = p * f2 + 1.0 / 15.0;
  p = p * f2 + 1.0 / 13.0;
  p = p * f2 + 1.0 / 11.0;
  p = p * f2 + 1.0 / 9.0;
  p = p * f2 + 1.0 / 7.0;
  p = p * f2 + 1.0 / 5.0;
  p = p * f2 + 1.0 / 3.0;
  p = p * f2 + 1.0;
  f64 lm = (2.0 * f) * p;
  const f64 ln2_hi = 6.93147180369123816490e-01;
  const f64 ln2_lo = 1.90821492927058770002e-10;
  f64 de = (f64)e;
  return de * ln2_hi + (lm + de * ln2_lo);
}

// Square root for x >= 1 by Newton's iteration from an exponent-halving first guess.
LDA_HD f64 det_sqrt(f64 x) {
  f64 y = f64_from_bits((f64_bits(x) >> 1) + 0x1FF8000000000000ull);
#pragma unroll
  for (int i = 0; i < 6; ++i) y = 0.5 * (y + x / y);
  return y;
}

// log(k!) for k >= 0: exact sum of logs below 16, Stirling's series (to x^-5) above.
LDA_HD f64 log_factorial(i64 k) {
  if (k < 16) {
    f64 s = 0.0;
    for (i64 i = 2; i <= k; ++i) s = s + det_log((f64)i);
    return s;
  }
  f64 x = (f64)(k + 1);
  f64 ix = 1.0 / x;
  f64 ix2 = ix * ix;
  f64 corr = ix * (1.0 / 12.0 - ix2 * (1.0 / 360.0 - ix2 * (1.0 / 1260.0)));
  const f64 half_log_2pi = 0.91893853320467274178;
  return (x - 0.5) * det_log(x) - x + half_log_2pi + corr;
}

// ---------------------------------------------------------------------------------------------
// Poisson draws for the Polya-urn step.  n < kPoisTables: alias table n (lambda = n + beta,
// values 0..kPoisValues-1), one Philox block.  n >= kPoisTables: Hoermann's PTRS
// ("The transformed rejection method for generating Poisson random variables", Insurance:
// Mathematics and Economics 12, 1993), one Philox block per attempt.
// ---------------------------------------------------------------------------------------------
constexpr u32 kPoisTables = 100;   // GPULDA_POIS_MAX_LAMBDA (src/cu/tuning.cuh:3)
constexpr u32 kPoisValues = 200;   // GPULDA_POIS_MAX_VALUE  (src/cu/tuning.cuh:4)
constexpr u32 kPtrsMaxAttempts = 256;

LDA_HD u32 poisson_ptrs(f64 lam, u32 c0, u32 c1, u32 c2, u32 tag, u32 k0, u32 k1) {
  f64 slam = det_sqrt(lam);
  f64 loglam = det_log(lam);
  f64 b = 0.931 + 2.53 * slam;
  f64 a = -0.059 + 0.02483 * b;
  f64 inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
  f64 vr = 0.9277 - 3.6224 / (b - 2.0);
  for (u32 att = 0; att < kPtrsMaxAttempts; ++att) {
    u32x4 r = philox_tagged(c0, c1, c2, tag, att, k0, k1);
    f64 U = u53(r.x, r.y) - 0.5;
    f64 V = u53(r.z, r.w);
    f64 us = 0.5 - (U < 0.0 ? -U : U);
    if (us <= 0.0) continue;
    f64 kf = (2.0 * a / us + b) * U + lam + 0.43;
    if (kf < 0.0 || kf >= 4294967295.0) continue;
    i64 k = (i64)kf;  // floor, kf >= 0
    if (us >= 0.07 && V <= vr) return (u32)k;
    if (us < 0.013 && V > us) continue;
    if (V <= 0.0) return (u32)k;  // log(0) = -inf accepts
    f64 lhs = det_log(V) + det_log(inv_alpha) - det_log(a / (us * us) + b);
    f64 rhs = -lam + (f64)k * loglam - log_factorial(k);
    if (lhs <= rhs) return (u32)k;
  }
  // Unreachable in practice (each attempt accepts with probability > 0.85).
  return (u32)(lam + 0.5);
}

// tables: kPoisTables rows of kPoisValues AliasSlots.
LDA_HD u32 poisson_draw(u32 n, f32 beta, const AliasSlot* tables, u32 c0, u32 c1, u32 c2, u32 tag,
                        u32 k0, u32 k1) {
  if (n < kPoisTables) {
    u32x4 r = philox_tagged(c0, c1, c2, tag, 0, k0, k1);
    return draw_alias_slot(tables + (u64)n * kPoisValues, kPoisValues, r.x, r.y);
  }
  f64 lam = (f64)n + (f64)beta;
  return poisson_ptrs(lam, c0, c1, c2, tag, k0, k1);
}

// ---------------------------------------------------------------------------------------------
// Exact alias split in fixed point (see the alias kernel): q = floor(w / total * n * 2^32).
// ---------------------------------------------------------------------------------------------

LDA_HD u64 alias_fixed_weight(f32 w, f32 total, u32 n) {
  f64 ratio = (f64)w / (f64)total;
  f64 q = ratio * ((f64)n * 4294967296.0);
  return (u64)q;
}

}  // namespace lda
