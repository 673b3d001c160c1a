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

__device__ __forceinline__ f32 readlane_f32(f32 v, u32 l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)l));
}
__device__ __forceinline__ u32 readlane_u32(u32 v, u32 l) {
  return (u32)__builtin_amdgcn_readlane((int)v, (int)l);
}

// Canonical inclusive f32 scan over the 64 lanes of a wavefront:
//   1. Hillis-Steele inside each 16-lane row with offsets 1, 2, 4, 8 (lanes below the offset
//      keep their value);
//   2. row carries c1 = T0, c2 = T0 + T1, c3 = c2 + T2 (T_r = total of row r), added to rows
//      1..3.
// The CPU checker evaluates the same network, so results agree bit for bit.
// Row shift-right by n inside 16-lane rows; lanes below n read 0 (DPP row_shr:n, bound_ctrl).
// Adding that 0 leaves the value unchanged, which is the "keep" of the row-local step; values
// here are never -0.0, so x + 0 == x bitwise.
template <int N>
__device__ __forceinline__ f32 row_shr_zero(f32 v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x110 + N, 0xF, 0xF, true));
}

__device__ __forceinline__ f32 wave_scan_f32(f32 v) {
  const u32 lane = lane_id();
  v = v + row_shr_zero<1>(v);
  v = v + row_shr_zero<2>(v);
  v = v + row_shr_zero<4>(v);
  v = v + row_shr_zero<8>(v);
  f32 t0 = readlane_f32(v, 15);
  f32 t1 = readlane_f32(v, 31);
  f32 t2 = readlane_f32(v, 47);
  f32 c2 = t0 + t1;
  f32 c3 = c2 + t2;
  f32 c = lane < 32 ? t0 : (lane < 48 ? c2 : c3);
  if (lane >= 16) v = v + c;
  return v;
}

// Inclusive u64 scan (exact, so any association is fine).
__device__ __forceinline__ u64 wave_scan_u64(u64 v) {
  const u32 lane = lane_id();
#pragma unroll
  for (u32 off = 1; off < 64; off <<= 1) {
    u64 t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// ---------------------------------------------------------------------------------------------
// Doc-topic map operations.  T points at `cap` LDS slots (cap a multiple of 64, never full).
// All 64 lanes of the wave call these with the same (uniform) arguments.
// ---------------------------------------------------------------------------------------------
struct MapProbe {
  u32 slot;   // slot of the key, or the slot where it belongs
  u32 found;  // 1 if present
  u32 entry;  // entry at `slot` when found
};

__device__ __forceinline__ MapProbe map_probe(const u32* T, u32 cap, u32 key) {
  const u32 lane = lane_id();
  const u32 h = map_home(key, cap);
  wave_sync();
  MapProbe p;
  p.slot = 0;
  p.found = 0;
  p.entry = kMapEmpty;
  for (u32 base = 0; base < cap; base += kMapLine) {
    u32 d = base + lane;
    u32 s = h + d;
    if (s >= cap) s -= cap;
    u32 e = T[s];
    u32 ek = map_key(e);
    bool empty = e == kMapEmpty;
    u32 de = empty ? 0 : map_disp(s, map_home(ek, cap), cap);
    bool match = !empty && ek == key;
    bool stop = empty || de < d || (de == d && ek > key);
    u64 bm = __ballot(match);
    u64 bs = __ballot(stop);
    u64 ev = bm | bs;
    if (ev) {
      u32 fl = first_lane(ev);
      u32 sl = h + base + fl;
      p.slot = sl >= cap ? sl - cap : sl;
      p.found = (u32)((bm >> fl) & 1ull);
      p.entry = readlane_u32(e, fl);
      return p;
    }
  }
  return p;
}

// Slot and entry of a key that is in the map.  Only the match is tested: a present key lies before
// the first empty slot of its probe sequence, so no stop test is needed.  (The empty entry's key
// field, 0xFFFF, never equals a topic.)
__device__ __forceinline__ MapProbe map_find(const u32* T, u32 cap, u32 key) {
  const u32 lane = lane_id();
  const u32 h = map_home(key, cap);
  wave_sync();
  MapProbe p;
  p.slot = 0;
  p.found = 0;
  p.entry = kMapEmpty;
  for (u32 base = 0; base < cap; base += kMapLine) {
    u32 s = h + base + lane;
    if (s >= cap) s -= cap;
    const u32 e = T[s];
    const u64 bm = __ballot(map_key(e) == key);
    if (bm) {
      const u32 fl = first_lane(bm);
      const u32 sl = h + base + fl;
      p.slot = sl >= cap ? sl - cap : sl;
      p.found = 1;
      p.entry = readlane_u32(e, fl);
      return p;
    }
  }
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
    u32 s = p.slot + base + lane;
    if (s >= cap) s -= cap;
    u32 e = T[s];
    u64 be = __ballot(e == kMapEmpty);
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
__device__ __forceinline__ u32 map_inc(u32* T, u32 cap, u32 key) { return map_inc_at(T, cap, key, map_probe(T, cap, key)); }

// Count of `key` -1 (the key must be present); deletes it by backward shift at zero.
// Returns 1 if the key was removed.
__device__ __forceinline__ u32 map_dec(u32* T, u32 cap, u32 key) {
  const u32 lane = lane_id();
  MapProbe p = map_find(T, cap, key);
  if (!p.found) return 0;  // not reached: every decremented topic is in the map
  if (map_count(p.entry) > 1) {
    if (lane == 0) T[p.slot] = p.entry - 1;
    return 0;
  }
  u32 start = p.slot + 1;
  if (start >= cap) start -= cap;
  for (u32 base = 0; base < cap; base += kMapLine) {
    wave_sync();
    u32 s = start + base + lane;
    if (s >= cap) s -= cap;
    u32 e = T[s];
    bool empty = e == kMapEmpty;
    bool stop = empty || map_disp(s, map_home(map_key(e), cap), cap) == 0;
    u64 bs = __ballot(stop);
    u32 fs = bs ? first_lane(bs) : 64;
    u32 dst = s == 0 ? cap - 1 : s - 1;
    if (lane <= fs) T[dst] = lane < fs ? e : kMapEmpty;
    if (bs) break;
  }
  return 1;
}

}  // namespace lda
