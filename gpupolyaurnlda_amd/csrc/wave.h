// wave.h -- wave64 building blocks for gfx950: canonical scan network, integer scans and the
// wave-cooperative Robin Hood doc-topic map in LDS.
//
// The doc-topic map replaces the half-warp HashMap of src/cu/hashmap.cuh:13-334 (insert2 /
// try_insert2 / get2 at :213-333).  One wavefront probes a whole 64-slot line with one LDS read
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

// ---------------------------------------------------------------------------------------------
// Doc-topic map operations.  T points at `cap` LDS slots (cap a multiple of 64, never full); hm is
// map_home_mult(K) for keys below K.  A slot is empty when its count is 0 (it then holds
// kMapEmpty).  All 64 lanes of the wave call these with the same (uniform) arguments.
// ---------------------------------------------------------------------------------------------
// x mod cap for x < 2 * cap: x - cap wraps past 2^32 when x < cap, so the minimum is the slot.
__device__ __forceinline__ u32 wrap_slot(u32 x, u32 cap) { return min(x, x - cap); }

struct MapProbe {
  u32 slot;       // slot of the key, or the slot where it belongs
  u32 found;      // 1 if present
  u32 entry;      // entry at `slot` when found
  u32 lane;       // lane of `slot` in the line the probe stopped in
  u32 line_slot;  // per lane: the slot this lane read in that line
  u32 line_e;     // per lane: the entry it read there
};

// Rewrites a line the wave has just read, with one lane's entry replaced: every lane stores its
// own slot, so no lane needs to be masked off (no exec-mask change on the scalar unit).
__device__ __forceinline__ void map_store_line(u32* T, u32 line_slot, u32 line_e, u32 at_lane, u32 entry) {
  T[line_slot] = lane_id() == at_lane ? entry : line_e;
}

__device__ __forceinline__ MapProbe map_probe(const u32* T, u32 cap, u32 hm, u32 key) {
  const u32 lane = lane_id();
  const u32 h = map_home(key, cap, hm);
  wave_sync();
  MapProbe p;
  p.slot = 0;
  p.found = 0;
  p.entry = kMapEmpty;
  p.lane = 0;
  p.line_slot = 0;
  p.line_e = kMapEmpty;
  for (u32 base = 0; base < cap; base += kMapLine) {
    u32 d = base + lane;
    const u32 s = wrap_slot(h + d, cap);
    u32 e = T[s];
    u32 ek = map_key(e);
    bool empty = map_is_empty(e);
    // (computed for empty slots too, where the stop mask below already holds)
    u32 de = map_disp(s, map_home(ek, cap, hm), cap);
    // match = !empty && ek == key; stop = empty || de < d || (de == d && ek > key), combined as
    // lane masks of single comparisons
    const u64 b_empty = ballot(empty);
    u64 bm = ballot(ek == key) & ~b_empty;
    u64 bs = b_empty | ballot(de < d) | (ballot(de == d) & ballot(ek > key));
    u64 ev = bm | bs;
    if (ev) {
      u32 fl = first_lane(ev);
      u32 sl = h + base + fl;
      p.slot = sl >= cap ? sl - cap : sl;
      p.found = (u32)((bm >> fl) & 1ull);
      p.entry = readlane_u32(e, fl);
      p.lane = fl;
      p.line_slot = s;
      p.line_e = e;
      return p;
    }
  }
  return p;
}

// Slot and entry of a key that is in the map.  Only the match is tested: a present key lies before
// the first empty slot of its probe sequence, so no stop test is needed.
__device__ __forceinline__ MapProbe map_find(const u32* T, u32 cap, u32 hm, u32 key) {
  const u32 lane = lane_id();
  const u32 h = map_home(key, cap, hm);
  wave_sync();
  MapProbe p;
  p.slot = 0;
  p.found = 0;
  p.entry = kMapEmpty;
  p.lane = 0;
  p.line_slot = 0;
  p.line_e = kMapEmpty;
  // the first line holds the key unless it was displaced past it (the common case is peeled)
  u32 base = 0;
  u32 e = T[wrap_slot(h + lane, cap)];
  u64 bm = ballot(!map_is_empty(e)) & ballot(map_key(e) == key);
  while (!bm) {
    base += kMapLine;
    if (base >= cap) return p;
    e = T[wrap_slot(h + base + lane, cap)];
    bm = ballot(!map_is_empty(e)) & ballot(map_key(e) == key);
  }
  const u32 fl = first_lane(bm);
  p.slot = wrap_slot(h + base + fl, cap);
  p.found = 1;
  p.entry = readlane_u32(e, fl);
  p.lane = fl;
  p.line_slot = wrap_slot(h + base + lane, cap);
  p.line_e = e;
  return p;
}

// Count of `key` +1 given its probe p; inserts it (Robin Hood shift) when absent.
__device__ __forceinline__ u32 map_inc_at(u32* T, u32 cap, u32 key, MapProbe p) {
  const u32 lane = lane_id();
  if (p.found) {
    if (lane == 0) T[p.slot] = p.entry + 1;
    return 0;
  }
  u32 carry = map_entry(key, 1);
  for (u32 base = 0; base < cap; base += kMapLine) {
    wave_sync();
    const u32 s = wrap_slot(p.slot + base + lane, cap);
    u32 e = T[s];
    u64 be = ballot(map_is_empty(e));
    u32 fe = be ? first_lane(be) : 64;
    u32 prev = __shfl_up(e, 1, 64);
    if (lane == 0) prev = carry;
    if (lane <= fe) T[s] = prev;
    if (be) break;
    carry = readlane_u32(e, 63);
  }
  return 1;
}

// Count of `key` +1; inserts it (Robin Hood shift) when absent.  Returns 1 if a key was added.
__device__ __forceinline__ u32 map_inc(u32* T, u32 cap, u32 hm, u32 key) {
  return map_inc_at(T, cap, key, map_probe(T, cap, hm, key));
}

// Count of `key` -1 (the key must be present); deletes it by backward shift at zero.
// Returns 1 if the key was removed.
__device__ __forceinline__ u32 map_dec(u32* T, u32 cap, u32 hm, u32 key) {
  const u32 lane = lane_id();
  MapProbe p = map_find(T, cap, hm, key);
  if (!p.found) return 0;  // not reached: every decremented topic is in the map
  if (map_count(p.entry) > 1) {
    map_store_line(T, p.line_slot, p.line_e, p.lane, p.entry - 1);
    return 0;
  }
  u32 start = p.slot + 1;
  if (start >= cap) start -= cap;
  for (u32 base = 0; base < cap; base += kMapLine) {
    wave_sync();
    const u32 s = wrap_slot(start + base + lane, cap);
    u32 e = T[s];
    u64 bs = ballot(map_is_empty(e)) | ballot(map_disp(s, map_home(map_key(e), cap, hm), cap) == 0);
    u32 fs = bs ? first_lane(bs) : 64;
    u32 dst = s == 0 ? cap - 1 : s - 1;
    if (lane <= fs) T[dst] = lane < fs ? e : kMapEmpty;
    if (bs) break;
  }
  return 1;
}

}  // namespace lda
