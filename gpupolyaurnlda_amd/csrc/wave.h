// wave.h -- wave64 building blocks for gfx950: canonical scan network, integer scans and the
// wave-cooperative doc-topic state in LDS.
//
// The doc-topic state replaces the half-warp HashMap of src/cu/hashmap.cuh:13-334 (insert2 /
// try_insert2 / get2 at :213-333).  It is split in two: a 16-bit count per topic, which a token
// changes in O(1) without a search, and a Robin Hood key table of the topics present, which the
// sparse sum walks 64 slots per step.  One wavefront probes a whole 64-slot line with one LDS read
// and a ballot, insertion shifts the tail of the run right by one slot (Robin Hood order is
// "sorted by home slot, then key"), and deletion shifts it back (the reference never deletes
// zero-count keys, so its table only grows and needs device-malloc resizes).  Because the order
// is total, the layout depends on the key set only; any insertion order gives the same table.
#pragma once

#include "lda_core.h"

namespace lda {

__device__ __forceinline__ u32 lane_id() { return __lane_id(); }

// Orders this lane's earlier LDS writes before later LDS reads of other lanes of the same
// wavefront (lanes of one wave exchange data through the doc-topic map and histograms).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Values the whole wave agrees on, moved to scalar registers: loops and branches on them are
// then uniform (no exec-mask bookkeeping) and their arithmetic runs on the scalar unit.
__device__ __forceinline__ u32 uniform_u32(u32 v) { return (u32)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ u64 uniform_u64(u64 v) {
  return (u64)uniform_u32((u32)v) | ((u64)uniform_u32((u32)(v >> 32)) << 32);
}

__device__ __forceinline__ u64 lanemask_lt() {
  u32 l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ u32 first_lane(u64 mask) { return (u32)(__ffsll((unsigned long long)mask) - 1); }

// Mask of the active lanes where pred holds.  The builtin takes the comparison's lane mask as it
// is (no 0/1 value materialised and compared again, as __ballot's int argument requires).
__device__ __forceinline__ u64 ballot(bool pred) { return (u64)__builtin_amdgcn_ballot_w64(pred); }

__device__ __forceinline__ f32 readlane_f32(f32 v, u32 l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)l));
}
__device__ __forceinline__ u32 readlane_u32(u32 v, u32 l) {
  return (u32)__builtin_amdgcn_readlane((int)v, (int)l);
}

// Canonical inclusive f32 scan over the 64 lanes of a wavefront:
//   1. Hillis-Steele inside each 16-lane row with offsets 1, 2, 4, 8 (lanes below the offset
//      keep their value);
//   2. rows 1 and 3 add the last value of rows 0 and 2 (DPP row_bcast:15);
//   3. rows 2 and 3 add lane 31 (row 1's last value after step 2: T1 + T0; DPP row_bcast:31).
//   So with row totals T_r: row 1 += T0, row 2 += T1 + T0, row 3 gets (x + T2) + (T1 + T0).
// The CPU checker evaluates the same network, so results agree bit for bit.
// Row shift-right by n inside 16-lane rows; lanes below n read 0 (DPP row_shr:n, bound_ctrl).
// Adding that 0 leaves the value unchanged, which is the "keep" of the row-local step; values
// here are never -0.0, so x + 0 == x bitwise.
template <int N>
__device__ __forceinline__ f32 row_shr_zero(f32 v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x110 + N, 0xF, 0xF, true));
}

// Lane l receives lane l - 1's value, lane 0 receives 0 (DPP wave_shr:1): an inclusive scan shifted
// into the exclusive one.
__device__ __forceinline__ f32 wave_shr1_zero(f32 v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, false));
}

// Broadcast steps: rows outside row_mask read 0 (and so keep their value after the add).
__device__ __forceinline__ f32 row_bcast15_odd(f32 v) {  // rows 1, 3 <- lane 15 of rows 0, 2
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0xA, 0xF, false));
}
__device__ __forceinline__ f32 row_bcast31_high(f32 v) {  // rows 2, 3 <- lane 31
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x143, 0xC, 0xF, false));
}

__device__ __forceinline__ f32 wave_scan_f32(f32 v) {
  v = v + row_shr_zero<1>(v);
  v = v + row_shr_zero<2>(v);
  v = v + row_shr_zero<4>(v);
  v = v + row_shr_zero<8>(v);
  v = v + row_bcast15_odd(v);
  v = v + row_bcast31_high(v);
  return v;
}

// Inclusive u64 scan over the same DPP network (integer sums are exact, so the network's
// association gives the plain prefix sums): each step moves both 32-bit halves with the same DPP
// control and adds them as one 64-bit value.
template <int CTRL, int ROW_MASK, bool BOUND_ZERO>
__device__ __forceinline__ u64 dpp_u64(u64 v) {
  const u32 lo = (u32)__builtin_amdgcn_update_dpp(0, (int)(u32)v, CTRL, ROW_MASK, 0xF, BOUND_ZERO);
  const u32 hi = (u32)__builtin_amdgcn_update_dpp(0, (int)(u32)(v >> 32), CTRL, ROW_MASK, 0xF, BOUND_ZERO);
  return ((u64)hi << 32) | lo;
}

__device__ __forceinline__ u64 wave_scan_u64(u64 v) {
  v += dpp_u64<0x111, 0xF, true>(v);  // row_shr:1
  v += dpp_u64<0x112, 0xF, true>(v);  // row_shr:2
  v += dpp_u64<0x114, 0xF, true>(v);  // row_shr:4
  v += dpp_u64<0x118, 0xF, true>(v);  // row_shr:8
  v += dpp_u64<0x142, 0xA, false>(v);  // row_bcast:15 into rows 1, 3
  v += dpp_u64<0x143, 0xC, false>(v);  // row_bcast:31 into rows 2, 3
  return v;
}

// Inclusive i32 max scan over the same DPP network (max is associative, so any network gives the
// plain prefix maxima): lanes outside a step's source keep their value (INT_MIN moved in).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_or_min(int v) {
  return __builtin_amdgcn_update_dpp((int)0x80000000, v, CTRL, ROW_MASK, 0xF, false);
}

__device__ __forceinline__ int wave_scan_max_i32(int v) {
  v = max(v, dpp_or_min<0x111, 0xF>(v));  // row_shr:1
  v = max(v, dpp_or_min<0x112, 0xF>(v));  // row_shr:2
  v = max(v, dpp_or_min<0x114, 0xF>(v));  // row_shr:4
  v = max(v, dpp_or_min<0x118, 0xF>(v));  // row_shr:8
  v = max(v, dpp_or_min<0x142, 0xA>(v));  // row_bcast:15 into rows 1, 3
  v = max(v, dpp_or_min<0x143, 0xC>(v));  // row_bcast:31 into rows 2, 3
  return v;
}

// Number of set lanes of `mask` below this lane.
__device__ __forceinline__ u32 lanes_below(u64 mask) {
  return __builtin_amdgcn_mbcnt_hi((u32)(mask >> 32), __builtin_amdgcn_mbcnt_lo((u32)mask, 0u));
}

// ---------------------------------------------------------------------------------------------
// Doc-topic state of one document in LDS: a count per topic and a Robin Hood key table.
//
//   C  (u16 per topic, K + 1 entries): n_dk; entry K is always 0.
//   Kt (u16 per slot, cap slots, cap a multiple of 64): 2 * topic for an occupied slot, 2 * K for an
//      empty one.  A slot's value is the byte offset of its count in C, and half the byte offset of
//      its phi value in the word's row; an empty slot thus reads the zero count C[K] and a phi
//      offset past the row (a buffer read that returns 0 without touching memory), so its mass is
//      0 * 0 = +0 with no test.
//
// The key table holds exactly the topics with a nonzero count, in the Robin Hood layout of linear
// probing with ties broken by key (deletion by backward shift).  That layout is a function of the
// key set and the capacity only.  With order-preserving homes (home non-decreasing in the key) it
// is the ordered layout: keys ascending, key i at slot (i + max(w, max_{j<=i}(h_j - j))) mod cap,
// h_j the home of key j and w the number of slots that the run passing the table's end occupies at
// its start.  Counts change by one per token through C alone, so a sweep touches the key table
// only when a topic enters or leaves the document.
//
// All 64 lanes call these functions with the same (uniform) arguments.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ u32 count_at(const u16* C, u32 key2) {
  return *(const u16*)((const char*)C + key2);
}
__device__ __forceinline__ void count_put(u16* C, u32 key2, u32 v) { *(u16*)((char*)C + key2) = (u16)v; }

// x mod cap for x < 2 * cap: x - cap wraps past 2^32 when x < cap, so the minimum is the slot.
__device__ __forceinline__ u32 wrap_slot(u32 x, u32 cap) { return min(x, x - cap); }

__device__ __forceinline__ u32 key_home(u32 key2, u32 cap, u32 hm) { return map_home(key2 >> 1, cap, hm); }

// Zeroes the K + 1 counts and counts the topics z[b .. b + L); returns the number of distinct
// topics (the lane whose atomic add finds a zero count is the first of its topic).  Counts are
// 16-bit halves of 32-bit words; a document has at most 65535 tokens, so no half carries.
__device__ __forceinline__ u32 doc_count_topics(u16* C, u32 K, const u16* __restrict__ z, u64 b, u32 L) {
  const u32 lane = lane_id();
  u32* Cw = (u32*)C;
  const u32 words = (K + 2) / 2;
  for (u32 i = lane; i < words; i += 64) Cw[i] = 0u;
  wave_sync();
  u32 fresh = 0;
  for (u32 i = lane; i < L; i += 64) {
    const u32 t = z[b + i];
    const u32 sh = (t & 1u) << 4;
    const u32 old = atomicAdd(&Cw[t >> 1], 1u << sh);
    fresh += ((old >> sh) & 0xFFFFu) == 0u ? 1u : 0u;
  }
  for (u32 off = 32; off > 0; off >>= 1) fresh += __shfl_xor(fresh, off, 64);
  wave_sync();
  return uniform_u32(fresh);
}

// Key table at capacity cap of the topics with a nonzero count (fewer than cap), in the ordered
// layout.  Two passes over the counts: the first finds their number n and max_j(h_j - j), hence
// the last key's unwrapped slot n - 1 + max_j(h_j - j) and w; the second places every key.
__device__ __forceinline__ void keymap_build(u16* Kt, u32 cap, u32 hm, const u16* C, u32 K) {
  const u32 lane = lane_id();
  const u32 e2 = 2u * K;
  for (u32 s = lane; s < cap; s += 64) Kt[s] = (u16)e2;
  int run = (int)0x80000000;
  u32 base = 0;
  for (u32 k0 = 0; k0 < K; k0 += 64) {
    const u32 k = k0 + lane;
    const bool present = k < K && C[k] != 0;
    const u64 bal = ballot(present);
    if (!bal) continue;
    const u32 i = base + lanes_below(bal);
    const int v = present ? (int)map_home(k, cap, hm) - (int)i : (int)0x80000000;
    run = max(run, (int)readlane_u32((u32)wave_scan_max_i32(v), 63));
    base += (u32)__popcll(bal);
  }
  if (base == 0) {
    wave_sync();
    return;
  }
  const int last = (int)base - 1 + run;
  int carry = last + 1 - (int)cap > 0 ? last + 1 - (int)cap : 0;
  wave_sync();
  base = 0;
  for (u32 k0 = 0; k0 < K; k0 += 64) {
    const u32 k = k0 + lane;
    const bool present = k < K && C[k] != 0;
    const u64 bal = ballot(present);
    if (!bal) continue;
    const u32 i = base + lanes_below(bal);
    const int v = present ? (int)map_home(k, cap, hm) - (int)i : (int)0x80000000;
    const int sc = wave_scan_max_i32(v);
    const int M = max(carry, sc);
    const u32 slot = wrap_slot(i + (u32)M, cap);  // < cap while fewer than cap keys (the callers' rule)
    if (present && slot < cap) Kt[slot] = (u16)(2u * k);
    carry = max(carry, (int)readlane_u32((u32)sc, 63));
    base += (u32)__popcll(bal);
  }
  wave_sync();
}

// Slot of a key that is in the table (kNoKeySlot if it is not: never the case for the callers).
constexpr u32 kNoKeySlot = 0xFFFFFFFFu;
__device__ __forceinline__ u32 keymap_find(const u16* Kt, u32 cap, u32 hm, u32 key2) {
  const u32 lane = lane_id();
  const u32 h = key_home(key2, cap, hm);
  wave_sync();
  for (u32 base = 0; base < cap; base += kMapLine) {
    const u64 bm = ballot(Kt[wrap_slot(h + base + lane, cap)] == key2);
    if (bm) return wrap_slot(h + base + first_lane(bm), cap);
  }
  return kNoKeySlot;
}

// Removes a key that is in the table: the entries after it move back one slot, up to the first
// empty slot or entry at its home.
__device__ __forceinline__ void keymap_erase(u16* Kt, u32 cap, u32 hm, u32 e2, u32 key2) {
  const u32 lane = lane_id();
  const u32 at = keymap_find(Kt, cap, hm, key2);
  if (at == kNoKeySlot) return;
  const u32 start = wrap_slot(at + 1u, cap);
  for (u32 base = 0; base < cap; base += kMapLine) {
    wave_sync();
    const u32 s = wrap_slot(start + base + lane, cap);
    const u32 v = Kt[s];
    const u64 bs = ballot(v == e2) | ballot(key_home(v, cap, hm) == s);
    const u32 fs = bs ? first_lane(bs) : 64u;
    const u32 dst = s == 0 ? cap - 1u : s - 1u;
    if (lane <= fs) Kt[dst] = (u16)(lane < fs ? v : e2);
    if (bs) break;
  }
  wave_sync();
}

// Adds a key that is not in the table (the table has a free slot): it goes to the first slot of
// its probe sequence (distance d from its home) that is empty or holds an entry displaced less
// than d, or displaced d with a larger key; the entries from there to the first empty slot move
// up one slot.
__device__ __forceinline__ void keymap_insert(u16* Kt, u32 cap, u32 hm, u32 e2, u32 key2) {
  const u32 lane = lane_id();
  const u32 h = key_home(key2, cap, hm);
  wave_sync();
  u32 at = h;
  for (u32 base = 0; base < cap; base += kMapLine) {
    const u32 d = base + lane;
    const u32 s = wrap_slot(h + d, cap);
    const u32 v = Kt[s];
    const u32 de = map_disp(s, key_home(v, cap, hm), cap);  // (for an empty slot the stop holds anyway)
    const u64 bs = ballot(v == e2) | ballot(de < d) | (ballot(de == d) & ballot(v > key2));
    if (bs) {
      at = wrap_slot(h + base + first_lane(bs), cap);
      break;
    }
  }
  u32 carry = key2;
  for (u32 base = 0; base < cap; base += kMapLine) {
    wave_sync();
    const u32 s = wrap_slot(at + base + lane, cap);
    const u32 v = Kt[s];
    const u64 be = ballot(v == e2);
    const u32 fe = be ? first_lane(be) : 64u;
    u32 prev = (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, false);  // wave_shr:1
    if (lane == 0) prev = carry;
    if (lane <= fe) Kt[s] = (u16)prev;
    if (be) break;
    carry = readlane_u32(v, 63);
  }
  wave_sync();
}

}  // namespace lda
