// kernels.hip -- gfx950 kernels of one Polya-urn LDA iteration.
//
//   k_phi_draw      Poisson Polya-urn draw phi~_vk ~ Pois(n_vk + beta) and exact topic sums
//                   (replaces polya_urn_sample, src/cu/polyaurn.cu:87-123, and polya_urn_init,
//                   :61-79, whose non-integer Gaussian branch and float atomics are dropped)
//   k_alias_rows    per-row normalise + canonical row sum + sigma_a + exact alias split
//                   (replaces polya_urn_colsums, polyaurn.cu:144-178, build_alias,
//                   src/cu/spalias.cu:88-174, and the cuBLAS transpose, polyaurn.cu:127-131,
//                   which word-major storage makes unnecessary)
//   k_sweep         one wavefront per document: the z-sweep over an LDS Robin Hood map
//                   (replaces sample_topics, src/cu/topics.cu:36-105)
//   k_count         word-sorted n_wk reduction through per-wave LDS histograms
//                   (replaces the per-token atomicAdd of topics.cu:95 and polya_urn_reset,
//                   polyaurn.cu:133-141)
#include "kernels.h"
#include "wave.h"

namespace lda {

// ---------------------------------------------------------------------------------------------
// Small utility kernels (fills are kernels, not hipMemsetAsync, so that an iteration can be
// captured into a hipGraph and replayed).
// ---------------------------------------------------------------------------------------------
__global__ void k_fill_u32(u32* p, u64 n, u32 v) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) p[i] = v;
}

// Iteration counter kept on the device, so that a captured iteration can be replayed.
__global__ void k_bump_u32(u32* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] += 1u;
}

__global__ void k_fill_u64(unsigned long long* p, u64 n, unsigned long long v) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) p[i] = v;
}

// u32 topic indicators at the ABI -> u16 on the device; counts indicators >= K into *bad.
__global__ void k_z_narrow(const u32* in, u16* out, u64 n, u32 K, u32* bad) {
  u32 nbad = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    u32 z = in[i];
    if (z >= K) {
      ++nbad;
      z = 0;
    }
    out[i] = (u16)z;
  }
  if (nbad) atomicAdd(bad, nbad);
}

__global__ void k_z_widen(const u16* in, u32* out, u64 n) {
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) out[i] = in[i];
}

// Initial topic indicators z_i ~ Uniform{0..K-1} keyed by (token, doc).
__global__ void k_z_init(u16* z, const u64* doc_off, u32 n_docs, u32 doc_base, u32 K, u32 k0, u32 k1) {
  u32 doc = blockIdx.x;
  if (doc >= n_docs) return;
  u64 b = doc_off[doc];
  u32 L = (u32)(doc_off[doc + 1] - b);
  for (u32 i = threadIdx.x; i < L; i += blockDim.x) {
    u32x4 r = philox_tagged(i, doc_base + doc, 0, kTagInitZ, 0, k0, k1);
    z[b + i] = (u16)bounded(r.x, K);
  }
}

// ---------------------------------------------------------------------------------------------
// Phi step.  Thread = topic column k, block.y = a tile of words; topic sums are exact u64.
// ---------------------------------------------------------------------------------------------
__global__ void k_phi_draw(const u32* __restrict__ n, f32* __restrict__ phi, unsigned long long* topic_sum,
                           u32 V, u32 K, f32 beta, u32 iter_arg, const u32* iter_ptr, u32 k0, u32 k1,
                           const AliasSlot* __restrict__ tables, u32 rows_per_block, u16* __restrict__ counts,
                           u8* __restrict__ wide) {
  u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const u32 iter = iter_ptr ? *iter_ptr : iter_arg;
  u32 v0 = blockIdx.y * rows_per_block;
  u32 v1 = v0 + rows_per_block < V ? v0 + rows_per_block : V;
  u64 acc = 0;
  // kPhiBatch words at a time, so that their counts and then their table slots are loaded
  // together rather than one dependent pair of loads per word: the draw of every cell is
  // poisson_draw's (the table draw below n = kPoisTables, PTRS above it)
  constexpr u32 kPhiBatch = 16;
  auto put = [&](u32 v, u32 xb) {
    const u64 idx = (u64)v * K + k;
    phi[idx] = (f32)xb;
    if (counts) {  // the u16 count row, and its word marked when a count passes 16 bits
      counts[idx] = xb <= 0xFFFFu ? (u16)xb : (u16)0xFFFFu;
      if (xb > 0xFFFFu) wide[v] = 1;
    }
    acc += xb;
  };
  for (u32 vb = v0; vb < v1; vb += kPhiBatch) {
    u32 nv[kPhiBatch], x[kPhiBatch], rk[kPhiBatch];
    AliasSlot sl[kPhiBatch];
#pragma unroll
    for (u32 b = 0; b < kPhiBatch; ++b) nv[b] = vb + b < v1 ? n[(u64)(vb + b) * K + k] : 0u;
#pragma unroll
    for (u32 b = 0; b < kPhiBatch; ++b) {
      const u32x4 r = philox_tagged(k, vb + b, iter, kTagPhi, 0, k0, k1);
      x[b] = bounded(r.x, kPoisValues);
      rk[b] = r.y;
      const u32 tb = nv[b] < kPoisTables ? nv[b] : 0u;
      sl[b] = tables[(u64)tb * kPoisValues + x[b]];
    }
    bool ptrs = false;
#pragma unroll
    for (u32 b = 0; b < kPhiBatch; ++b) {
      if (vb + b < v1 && nv[b] < kPoisTables) put(vb + b, u24(rk[b]) < sl[b].prob ? x[b] : sl[b].alias);
      ptrs = ptrs || nv[b] >= kPoisTables;
    }
    if (ptrs) {  // the words whose count is at least kPoisTables, one at a time (their count loaded again)
#pragma unroll 1
      for (u32 v = vb; v < v1 && v < vb + kPhiBatch; ++v) {
        const u32 nb = n[(u64)v * K + k];
        if (nb >= kPoisTables) put(v, poisson_ptrs((f64)nb + (f64)beta, k, v, iter, kTagPhi, k0, k1));
      }
    }
  }
  if (acc) atomicAdd(&topic_sum[k], (unsigned long long)acc);
}

// n_vk ~ Pois(floor(C_v / K) + beta) at initialisation (polya_urn_init, polyaurn.cu:61-79).
__global__ void k_init_counts(const u32* __restrict__ C, u32* __restrict__ n, u32 V, u32 K, f32 beta, u32 k0,
                              u32 k1, const AliasSlot* __restrict__ tables, u32 rows_per_block) {
  u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  u32 v0 = blockIdx.y * rows_per_block;
  u32 v1 = v0 + rows_per_block < V ? v0 + rows_per_block : V;
  for (u32 v = v0; v < v1; ++v)
    n[(u64)v * K + k] = poisson_draw(C[v] / K, beta, tables, k, v, 0, kTagInitN, k0, k1);
}

// ---------------------------------------------------------------------------------------------
// Alias rows.  One wavefront per row of length m:
//   A. (optional) x_k /= (f32)divisor_k; canonical row sum T (lane l adds x_l, x_{64+l}, ... in
//      order, then the canonical wave scan); sigma = alpha * T.  The row is staged in LDS.
//   B. q_k = floor(x_k / T * m * 2^32) (exact double ops), slot capacity Q = sum_k q_k and item
//      weight a_k = m q_k, so that total deficit equals total surplus exactly.  a < Q is light
//      (deficit Q - a), otherwise heavy (surplus a - Q).  Exclusive deficit prefixes D over the
//      lights and inclusive surplus prefixes E over the heavies, in index order, are exact u64
//      sums (lane l takes slots [l*per, (l+1)*per) in order; one wave scan of the lane totals
//      places them); they share one LDS array of m entries (the nh heavies' E from the front,
//      the nl lights' D from the back, nh + nl = m), as do the slot indices.
//   C. light i: alias = first heavy with E > D_i, prob = a_i / Q.
//      heavy j: its turn ends at D_end = D of the first light with D >= E_j, or the total deficit
//      when there is none; if D_end > E_j it keeps prob = (Q + E_j - D_end) / Q and its
//      alias is the next heavy, otherwise (or with no next heavy) prob = 1, alias = j.
//      Both are ranks between the ascending sequences D' = D + 1 (lights) and E (heavies): in
//      their merge with ties taken from D' first, light i is preceded by exactly the heavies with
//      E < D_i + 1 and heavy j by exactly the lights with D + 1 <= E_j.  Each lane produces `per`
//      consecutive outputs of that merge (one diagonal binary search, then a sequential merge).
// This is Vose's pairing expressed through prefix sums (the sweep of Huebschle-Schneider &
// Sanders, "Parallel Weighted Random Sampling", 2019), so every slot is computed independently.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ u64 readlane_u64(u64 v, u32 l) {
  u32 lo = readlane_u32((u32)v, l);
  u32 hi = readlane_u32((u32)(v >> 32), l);
  return ((u64)hi << 32) | lo;
}

// a / b for a row's fixed divisor b, from y = 1 / b (correctly rounded, once per row): q = a y is
// within an ulp of a / b, the residual a - b q is exact in one FMA, and one FMA correction then
// gives the correctly rounded quotient (Markstein, IBM J. Res. Dev. 34(1), 1990, Theorem 1), the
// same bits as the division.  Operands here are non-negative and far from overflow and underflow.
__device__ __forceinline__ f64 div_by(f64 a, f64 b, f64 y) {
  const f64 q = a * y;
  const f64 r = __builtin_fma(-q, b, a);
  return __builtin_fma(r, y, q);
}

// alias_fixed_weight with the row total's reciprocal: floor(w / T * n * 2^32), which is below
// 2^52 (w <= T, n <= 2^20), converted to u64 by adding 2^52 (exact for integers below 2^52, whose
// bits then sit in the mantissa).
__device__ __forceinline__ u64 alias_fixed_weight_r(f32 w, f32 T, f64 yT, u32 n) {
  const f64 q = div_by((f64)w, (f64)T, yT) * ((f64)n * 4294967296.0);
  return (u64)__double_as_longlong(__builtin_trunc(q) + 0x1p52) - 0x4330000000000000ull;
}

// PER > 0: rows of at most 64 * PER slots, each lane's fixed-point weights held in registers;
// PER == 0: any row length, weights recomputed in each pass.
template <int PER>
__global__ void k_alias_rows(f32* __restrict__ rows, u32 n_rows, u32 m, const unsigned long long* __restrict__ divisor,
                             f32* __restrict__ sigma, f32 alpha, AliasSlot* __restrict__ out, u32 rows_per_block_waves) {
  extern __shared__ unsigned char alias_smem[];
  const u32 lane = lane_id();
  const u32 wave = threadIdx.x >> 6;
  const u32 row = blockIdx.x * rows_per_block_waves + wave;
  if (row >= n_rows) return;
  // per wavefront (alias_lds_per_wave): m4 u64 prefixes, m4 + 68 f32 row values (a light's
  // probability once its prefix is placed), m4 u16 slot indices.  Lane l works on slots
  // [l * per, (l + 1) * per); row values sit at xp(k) = k + pad * (k / per) with pad making the
  // lanes' stride per + pad odd, so the lanes' reads of their t-th slot fall in distinct banks.
  const u32 m4 = (m + 3) & ~3u;
  u64* pre = reinterpret_cast<u64*>(alias_smem + (u64)wave * alias_lds_per_wave(m));
  f32* xs = reinterpret_cast<f32*>(pre + m4);
  u16* idx = reinterpret_cast<u16*>(xs + m4 + 68);
  const u32 per = (m + 63) / 64;  // at most PER when PER > 0
  const u32 pad = (per & 1u) ^ 1u;
  // k / per as a multiply-high by ceil(2^32 / per): exact for k < 2^16 and per < 2^16 (the
  // reciprocal's excess adds less than k / 2^32 < 2^-16 < 1 / per to a quotient whose fraction
  // is at most 1 - 1 / per); per == 1 has pad 0, so its quotient is not used
  const u32 per_magic = per > 1 ? 0xFFFFFFFFu / per + 1u : 0u;  // ceil(2^32 / per), per <= 2^16
  auto xp = [&](u32 k) { return k + pad * __umulhi(k, per_magic); };

  f32* x = rows + (u64)row * m;
  AliasSlot* o = out + (u64)row * m;

  f32 acc = 0.0f;
  if constexpr (PER > 0) {
    // every load of the lane's PER strided slots in flight before any arithmetic
    f32 xv[PER];
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      const u32 k = (u32)c * 64 + lane;
      xv[c] = k < m ? x[k] : 0.0f;
    }
    if (divisor) {
      u64 dr[PER];  // raw loads first: converting each as it arrives would wait on it
#pragma unroll
      for (int c = 0; c < PER; ++c) {
        const u32 k = (u32)c * 64 + lane;
        dr[c] = k < m ? divisor[k] : 0ull;
      }
#pragma unroll
      for (int c = 0; c < PER; ++c) {
        const u32 k = (u32)c * 64 + lane;
        if (k < m) {
          const f32 d = (f32)(f64)dr[c];
          xv[c] = d > 0.0f ? xv[c] / d : 0.0f;
          x[k] = xv[c];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      const u32 k = (u32)c * 64 + lane;
      if (k < m) {
        xs[xp(k)] = xv[c];
        acc = acc + xv[c];
      }
    }
  } else {
    for (u32 c = 0; c < m; c += 64) {
      u32 k = c + lane;
      if (k < m) {
        f32 p = x[k];
        if (divisor) {
          f32 d = (f32)(f64)divisor[k];
          p = d > 0.0f ? p / d : 0.0f;
          x[k] = p;
        }
        xs[xp(k)] = p;
        acc = acc + p;
      }
    }
  }
  f32 T = readlane_f32(wave_scan_f32(acc), 63);
  if (sigma && lane == 0) sigma[row] = alpha * T;
  if (!(T > 0.0f)) {
    for (u32 k = lane; k < m; k += 64) {
      AliasSlot s;
      s.prob = 1.0f;
      s.alias = k;
      o[k] = s;
    }
    return;
  }
  wave_sync();
  const f64 yT = 1.0 / (f64)T;

  // B1: fixed-point weights and their exact total Q (the capacity of one slot)
  const u32 k0 = lane * per < m ? lane * per : m;
  const u32 k1 = k0 + per < m ? k0 + per : m;
  const u32 own = k0 + pad * lane;  // xp(k0): this lane's slots k0 + t sit at own + t
  u64 qr[PER > 0 ? PER : 1];
  u64 qsum = 0;
  if constexpr (PER > 0) {
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      qr[t] = (u32)t < per && k0 + t < k1 ? alias_fixed_weight_r(xs[own + t], T, yT, m) : 0;
      qsum += qr[t];
    }
  } else {
    for (u32 t = 0; k0 + t < k1; ++t) qsum += alias_fixed_weight_r(xs[own + t], T, yT, m);
  }
  const u64 Q = readlane_u64(wave_scan_u64(qsum), 63);
  const f64 yQ = 1.0 / (f64)Q;
  auto weight = [&](u32 t) -> u64 {  // a = m q of slot k0 + t
    if constexpr (PER > 0) return qr[t] * m;
    else return alias_fixed_weight_r(xs[own + t], T, yT, m) * m;
  };

  // B2: lane totals of surplus, deficit and class counts, placed by wave scans
  u64 sur = 0, def = 0;
  u32 nh_l = 0;
  constexpr u32 kUnroll = PER > 0 ? (u32)PER : 1u;
  const u32 steps = PER > 0 ? kUnroll : per;
#pragma unroll
  for (u32 t = 0; t < steps; ++t) {
    if (k0 + t < k1) {
      const u64 a = weight(t);
      if (a >= Q) {
        sur += a - Q;
        ++nh_l;
      } else {
        def += Q - a;
      }
    }
  }
  const u64 incS = wave_scan_u64(sur), incD = wave_scan_u64(def);
  const u64 incH = wave_scan_u64((u64)nh_l);
  const u32 nh = (u32)readlane_u64(incH, 63);
  const u32 nl = m - nh;
  const u64 sD = readlane_u64(incD, 63);
  // B3: E (inclusive, heavies from the front) and D (exclusive, lights from the back); a light's
  // probability a / Q replaces its row value
  {
    u64 E = incS - sur, D = incD - def;
    u32 h = (u32)(incH - nh_l), l = (k0 - h);  // lights before this lane = slots before it - heavies
#pragma unroll
    for (u32 t = 0; t < steps; ++t) {
      if (k0 + t < k1) {
        const u32 k = k0 + t;
        const u64 a = weight(t);
        if (a >= Q) {
          E += a - Q;
          pre[h] = E;
          idx[h] = (u16)k;
          ++h;
        } else {
          pre[m - 1 - l] = D;
          idx[m - 1 - l] = (u16)k;
          xs[own + t] = (f32)div_by((f64)a, (f64)Q, yQ);
          D += Q - a;
          ++l;
        }
      }
    }
  }
  wave_sync();

  // C: merge of D' = D + 1 (lights, A) and E (heavies, B), ties to A; this lane's outputs are
  // merge positions [k0, k1)
  u32 lo = k0 > nh ? k0 - nh : 0, hi = k0 < nl ? k0 : nl;
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if (pre[m - 1 - mid] + 1 <= pre[k0 - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  u32 i = lo, j = k0 - lo;
  for (u32 t = k0; t < k1; ++t) {
    const bool take_light = j >= nh || (i < nl && pre[m - 1 - i] + 1 <= pre[j]);
    AliasSlot s;
    u32 k;
    if (take_light) {
      // j heavies have E <= D_i: the first with E > D_i is heavy j (the last one if none)
      k = idx[m - 1 - i];
      s.prob = xs[xp(k)];
      s.alias = idx[j < nh ? j : nh - 1];
      ++i;
    } else {
      // i lights have D < E_j: the first with D >= E_j is light i
      k = idx[j];
      const u64 E = pre[j];
      const u64 dend = i < nl ? pre[m - 1 - i] : sD;
      s.prob = 1.0f;
      s.alias = k;
      if (dend > E && j + 1 < nh) {
        s.prob = (f32)div_by((f64)(Q + E - dend), (f64)Q, yQ);  // Q + E - dend < Q
        s.alias = idx[j + 1];
      }
      ++j;
    }
    o[k] = s;
  }
}
template __global__ void k_alias_rows<16>(f32*, u32, u32, const unsigned long long*, f32*, f32, AliasSlot*, u32);
template __global__ void k_alias_rows<0>(f32*, u32, u32, const unsigned long long*, f32*, f32, AliasSlot*, u32);

// Alias rows, one workgroup of kAliasBlockThreads (four wavefronts) per row: the same steps A-C
// with thread t working on slots [t * per, (t + 1) * per), per = ceil(m / 256) <= PER.  The row
// sum stays the canonical one (wavefront 0, lane l adding x_l, x_{64+l}, ... in order, then the
// canonical scan), and the integer sums and prefixes are exact, so every slot gets the bits of the
// one-wavefront kernel; each thread's share of the dependent steps (prefix placement, merge) is a
// quarter as long.
template <int PER>
__global__ __launch_bounds__(256) void k_alias_block(f32* __restrict__ rows, u32 n_rows, u32 m,
                                                    const unsigned long long* __restrict__ divisor,
                                                    f32* __restrict__ sigma, f32 alpha, AliasSlot* __restrict__ out,
                                                    u8* __restrict__ wide, u8* __restrict__ fit_out,
                                                    const u8* __restrict__ keep_f32,
                                                    const float2* __restrict__ recip) {
  extern __shared__ unsigned char alias_smem[];
  __shared__ unsigned long long red[3][kAliasBlockThreads / 64];
  __shared__ f32 shT;
  // rows last to first: the Poisson kernel wrote them first to last, so the rows still in the MALL
  // are read before later reads displace them
  if (blockIdx.x >= n_rows) return;
  const u32 row = n_rows - 1u - blockIdx.x;
  const u32 t = threadIdx.x;
  const u32 lane = lane_id();
  const u32 wave = t >> 6;
  const u32 m4 = (m + 3) & ~3u;
  u64* pre = reinterpret_cast<u64*>(alias_smem);
  f32* xs = reinterpret_cast<f32*>(pre + m4);
  u16* idx = reinterpret_cast<u16*>(xs + m4 + kAliasBlockThreads + 4);
  f32* lp = xs;  // after B1 the row values are dead: lights' probabilities, indexed like pre
  const u32 per = (m + kAliasBlockThreads - 1) / kAliasBlockThreads;  // <= PER
  const u32 pad = (per & 1u) ^ 1u;  // thread stride per + pad odd: distinct banks
  const u32 per_magic = per > 1 ? 0xFFFFFFFFu / per + 1u : 0u;  // ceil(2^32 / per), per <= 2^16
  auto xp = [&](u32 k) { return k + pad * __umulhi(k, per_magic); };
  f32* x = rows + (u64)row * m;
  AliasSlot* o = out + (u64)row * m;

  // whether the sweep is to read the row's u16 counts, read with the row by a thread off the canonical
  // sum's path (wavefront 1); the row's mark is cleared for the next step
  const bool fit_lane = fit_out && t == 64;
  u32 unfit = 0;

  // A: loads (coalesced), normalisation, the row staged in LDS; the canonical sum by wavefront 0
  {
    f32 xv[PER];
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      const u32 k = (u32)c * kAliasBlockThreads + t;
      xv[c] = k < m ? x[k] : 0.0f;
    }
    // a wavefront whose values are all counts below 2^16 divides through the divisors' reciprocals
    // (phi_from_count_f: the bits of the division, as the sweep's count rows recover them); the
    // reciprocals are loaded with the row
    const bool with_recip = divisor && recip;
    float2 rc[PER];
    if (with_recip) {
#pragma unroll
      for (int c = 0; c < PER; ++c) {
        const u32 k = (u32)c * kAliasBlockThreads + t;
        rc[c] = k < m ? recip[k] : make_float2(0.0f, 1.0f);
      }
    }
    if (fit_lane) {  // the flags, loaded behind the row and its reciprocals
      unfit = wide[row];
      if (keep_f32) unfit |= keep_f32[row];
    }
    bool by_recip = false;
    if (with_recip) {
      bool big = false;
#pragma unroll
      for (int c = 0; c < PER; ++c) big = big || xv[c] > 65535.0f;
      by_recip = ballot(big) == 0ull;
    }
    if (by_recip) {
#pragma unroll
      for (int c = 0; c < PER; ++c) {
        const u32 k = (u32)c * kAliasBlockThreads + t;
        if (k < m) {
          xv[c] = phi_from_count_f(xv[c], rc[c].x, rc[c].y);
          x[k] = xv[c];
        }
      }
    } else if (divisor) {
      u64 dr[PER];
#pragma unroll
      for (int c = 0; c < PER; ++c) {
        const u32 k = (u32)c * kAliasBlockThreads + t;
        dr[c] = k < m ? divisor[k] : 0ull;
      }
#pragma unroll
      for (int c = 0; c < PER; ++c) {
        const u32 k = (u32)c * kAliasBlockThreads + t;
        if (k < m) {
          const f32 d = (f32)(f64)dr[c];
          xv[c] = d > 0.0f ? xv[c] / d : 0.0f;
          x[k] = xv[c];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < PER; ++c) {
      const u32 k = (u32)c * kAliasBlockThreads + t;
      if (k < m) xs[xp(k)] = xv[c];
    }
  }
  __syncthreads();
  if (fit_lane) {
    fit_out[row] = unfit ? (u8)0 : (u8)1;
    wide[row] = 0;
  }
  if (wave == 0) {
    f32 acc = 0.0f;
    for (u32 k = lane; k < m; k += 64) acc = acc + xs[xp(k)];
    const f32 T0 = readlane_f32(wave_scan_f32(acc), 63);
    if (lane == 0) {
      shT = T0;
      if (sigma) sigma[row] = alpha * T0;
    }
  }
  __syncthreads();
  const f32 T = shT;
  if (!(T > 0.0f)) {
    for (u32 k = t; k < m; k += kAliasBlockThreads) {
      AliasSlot sl;
      sl.prob = 1.0f;
      sl.alias = k;
      o[k] = sl;
    }
    return;
  }
  const f64 yT = 1.0 / (f64)T;

  // B1: fixed-point weights and their exact total Q
  const u32 k0 = t * per < m ? t * per : m;
  const u32 k1 = k0 + per < m ? k0 + per : m;
  const u32 own = k0 + pad * t;  // xp(k0) for k0 < m
  u64 qr[PER];
  u64 qsum = 0;
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    qr[s] = (u32)s < per && k0 + s < k1 ? alias_fixed_weight_r(xs[own + s], T, yT, m) : 0;
    qsum += qr[s];
  }
  // B2: thread totals of surplus, deficit and heavies, with Q, placed by block scans
  const u64 wq = wave_scan_u64(qsum);
  if (lane == 63) red[0][wave] = wq;
  __syncthreads();
  u64 Q = 0;
#pragma unroll
  for (u32 w = 0; w < kAliasBlockThreads / 64; ++w) Q += red[0][w];
  const f64 yQ = 1.0 / (f64)Q;
  u64 sur = 0, def = 0;
  u32 nh_t = 0;
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    if ((u32)s < per && k0 + s < k1) {
      const u64 a = qr[s] * m;
      if (a >= Q) {
        sur += a - Q;
        ++nh_t;
      } else {
        def += Q - a;
      }
    }
  }
  const u64 wS = wave_scan_u64(sur), wD = wave_scan_u64(def), wH = wave_scan_u64((u64)nh_t);
  __syncthreads();  // everyone has read red[0]
  if (lane == 63) {
    red[0][wave] = wS;
    red[1][wave] = wD;
    red[2][wave] = wH;
  }
  __syncthreads();
  u64 oS = 0, oD = 0, oH = 0, sD = 0, nh64 = 0;
#pragma unroll
  for (u32 w = 0; w < kAliasBlockThreads / 64; ++w) {
    if (w < wave) {
      oS += red[0][w];
      oD += red[1][w];
      oH += red[2][w];
    }
    sD += red[1][w];
    nh64 += red[2][w];
  }
  const u32 nh = (u32)nh64;
  const u32 nl = m - nh;
  // B3: E (inclusive, heavies from the front) and D (exclusive, lights from the back); a light's
  // probability a / Q replaces its row value
  {
    u64 E = oS + wS - sur, D = oD + wD - def;
    u32 h = (u32)(oH + wH - nh_t), l = k0 - h;
#pragma unroll
    for (int s = 0; s < PER; ++s) {
      if ((u32)s < per && k0 + s < k1) {
        const u32 k = k0 + s;
        const u64 a = qr[s] * m;
        if (a >= Q) {
          E += a - Q;
          pre[h] = E;
          idx[h] = (u16)k;
          ++h;
        } else {
          pre[m - 1 - l] = D;
          idx[m - 1 - l] = (u16)k;
          lp[m - 1 - l] = (f32)div_by((f64)a, (f64)Q, yQ);
          D += Q - a;
          ++l;
        }
      }
    }
  }
  __syncthreads();

  // C: merge of D' = D + 1 (lights) and E (heavies), ties to the lights; this thread's outputs
  // are merge positions [k0, k1)
  u32 lo = k0 > nh ? k0 - nh : 0, hi = k0 < nl ? k0 : nl;
  while (lo < hi) {
    const u32 mid = (lo + hi) >> 1;
    if (pre[m - 1 - mid] + 1 <= pre[k0 - 1 - mid]) lo = mid + 1;
    else hi = mid;
  }
  // the two heads of the merge stay in registers: Dh = D of light i (sD past the last light),
  // Eh = E of heavy j (~0 past the last heavy, so a light is taken)
  u32 i = lo, j = k0 - lo;
  u64 Dh = i < nl ? pre[m - 1 - i] : sD;
  u64 Eh = j < nh ? pre[j] : ~0ull;
  for (u32 q = k0; q < k1; ++q) {
    AliasSlot sl;
    u32 k;
    if (i < nl && Dh + 1 <= Eh) {
      k = idx[m - 1 - i];
      sl.prob = lp[m - 1 - i];
      sl.alias = idx[j < nh ? j : nh - 1];
      ++i;
      Dh = i < nl ? pre[m - 1 - i] : sD;
    } else {
      k = idx[j];
      sl.prob = 1.0f;
      sl.alias = k;
      if (Dh > Eh && j + 1 < nh) {  // Dh: D of the first light with D >= E_j, or the total deficit
        sl.prob = (f32)div_by((f64)(Q + Eh - Dh), (f64)Q, yQ);
        sl.alias = idx[j + 1];
      }
      ++j;
      Eh = j < nh ? pre[j] : ~0ull;
    }
    o[k] = sl;
  }
}
template __global__ void k_alias_block<4>(f32*, u32, u32, const unsigned long long*, f32*, f32, AliasSlot*, u8*, u8*,
                                          const u8*, const float2*);
template __global__ void k_alias_block<16>(f32*, u32, u32, const unsigned long long*, f32*, f32, AliasSlot*, u8*, u8*,
                                           const u8*, const float2*);

// ---------------------------------------------------------------------------------------------
// z-sweep, one wavefront per document, the doc-topic state resident in LDS: a count per topic and
// the Robin Hood key table of the topics present (wave.h).
// Per token i (word w, old topic z), with r = Philox(i, doc, iter, SWEEP):
//   remove z (count - 1; the topic leaves the table at 0); if sigma_a[w] == 0 (phi_w. is all zero)
//   keep z; else
//   m[j][l] = count * phi[w][key] of table slot j*64 + l (line j, lane l; 0 when empty);
//   G_g = column sums of line group g = j % 4, lines left to right; A = ((G_0 + G_1) + G_2) + G_3;
//   P = canonical scan of A; sigma_b = P[63];
//   sparse if u(r.x) * (sigma_a + sigma_b) > sigma_a: target t = u(r.w) * sigma_b, lane l* = the
//   first with A > 0 and P > t (else the last with A > 0), E = P[l* - 1]; Q = canonical scan of
//   column l*'s masses in line order (64 lines per scan, carry = total of the earlier scans);
//   the line is the first with m > 0 and E + (carry + Q_j) > t (else the last with m > 0);
//   otherwise the alias table of w (slot from r.y, keep test on r.z); add the new topic (count + 1;
//   it enters the table at 1).
// Each lane sums its own column, so a token costs one wave scan however many lines the map has.
// A token's removal is made at the end of the token before it, in the same LDS instruction as that
// token's addition, so a token that changes no key set (most of them) finds its lines in registers
// and never searches the table.
// ---------------------------------------------------------------------------------------------
constexpr u32 kNoSlot = 0xFFFFFFFFu;

// The Phi row of a token's word as a buffer resource of K floats: a read at a topic past K (the key
// of an empty slot) returns 0 without touching memory, and the address path of a buffer load (a
// 32-bit offset under a scalar base) processes a wavefront's scattered reads faster than a global
// load does.
// A word's row of phi as the sweep reads it: the f32 row, or (counts == true) the row of phi~
// counts as u16 with the per-topic (RN(1 / S_k), S_k) in LDS, from which each value is recovered
// exactly.  Either way a buffer resource bounds the row, so the empty slot's offset (past the row)
// reads 0.
struct PhiRow {
  __amdgpu_buffer_rsrc_t rs;
  const float2* recip;  // LDS, K + 1 entries (counts rows only)
  bool counts;
};
constexpr int kBufferWord3 = 0x00020000;  // raw 32-bit buffer access (gfx9 family)

__device__ __forceinline__ PhiRow phi_row(const f32* prow, u32 K) {
  PhiRow r;
  r.rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<f32*>(prow), (short)0, (int)(K * sizeof(f32)), kBufferWord3);
  r.recip = nullptr;
  r.counts = false;
  return r;
}
__device__ __forceinline__ PhiRow phi_count_row(const u16* crow, u32 K, const float2* recip) {
  PhiRow r;
  r.rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<u16*>(crow), (short)0, (int)(K * sizeof(u16)), kBufferWord3);
  r.recip = recip;
  r.counts = true;
  return r;
}

// phi[w][topic] for a table value key2 = 2 * topic: the f32 row at byte offset 2 * key2, the u16
// count row at byte offset key2.
__device__ __forceinline__ f32 row_at2(const PhiRow& row, u32 key2) {
  return __builtin_bit_cast(f32, __builtin_amdgcn_raw_buffer_load_b32(row.rs, (int)(key2 * 2u), 0, 0));
}
__device__ __forceinline__ u32 count_row_at2(const PhiRow& row, u32 key2) {
  return (u32)__builtin_amdgcn_raw_buffer_load_b16(row.rs, (int)key2, 0, 0);
}
// The reciprocal pair of a table value key2 = 2 * topic (byte offset 4 * key2: one shift-add).
__device__ __forceinline__ float2 recip_at2(const PhiRow& row, u32 key2) {
  return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(row.recip) + key2 * 4u);
}
__device__ __forceinline__ f32 row_value(const PhiRow& row, u32 key2) {
  if (row.counts) {
    const float2 rc = recip_at2(row, key2);
    return phi_from_count(count_row_at2(row, key2), rc.x, rc.y);
  }
  return row_at2(row, key2);
}
// N values of a row, every load issued before any arithmetic on them.
template <int N>
__device__ __forceinline__ void row_values(const PhiRow& row, const u32 (&e)[N], f32 (&f)[N]) {
  if (row.counts) {
    u32 x[N];
#pragma unroll
    for (int j = 0; j < N; ++j) x[j] = count_row_at2(row, e[j]);
    // the reciprocals (LDS) four at a time, which keeps the registers of eight lines in budget
#pragma unroll
    for (int j0 = 0; j0 < N; j0 += 4) {
      float2 rc[4];
#pragma unroll
      for (int j = j0; j < j0 + 4 && j < N; ++j) rc[j - j0] = recip_at2(row, e[j]);
#pragma unroll
      for (int j = j0; j < j0 + 4 && j < N; ++j) f[j] = phi_from_count(x[j], rc[j - j0].x, rc[j - j0].y);
    }
  } else {
#pragma unroll
    for (int j = 0; j < N; ++j) f[j] = row_at2(row, e[j]);
  }
}

// Mass of one table slot, count * phi: an empty slot (2K) reads the zero count C[K] and a phi offset
// past the row, so its mass is 0 * 0 = +0, the value the definition gives it, with no branch (all
// reads of a token in flight together).
__device__ __forceinline__ f32 slot_mass(const u16* C, u32 key2, const PhiRow& row) {
  return (f32)count_at(C, key2) * row_value(row, key2);
}

__device__ __forceinline__ f32 combine_groups(f32 g0, f32 g1, f32 g2, f32 g3) { return ((g0 + g1) + g2) + g3; }

// The LDS of one document: counts at the start, the key table after them.
struct DocLds {
  u16* C;
  u16* Kt;
  const float2* recip;  // the block's copy of topic_recip (sweeps that read count rows), else null
};
__device__ __forceinline__ DocLds doc_lds(u32* smem, u32 K, const float2* recip = nullptr) {
  DocLds d;
  d.C = (u16*)smem;
  d.Kt = (u16*)((char*)smem + doc_counts_bytes(K));
  d.recip = recip;
  return d;
}

// Per-document state of a sweep.
struct DocState {
  u64 b;          // first token
  u32 L;          // tokens
  u32 cap_max;    // capacity for min(L, K) topics: the map never grows beyond it
  u32 kmax;       // min(L, K)
  u32 cap;        // current capacity
  u32 occupied;   // distinct topics in the map
  u32 gdoc;       // global document id (Philox counter word 1)
  u32 hm;         // map_home_mult(K)
  u32 e2;         // table value of an empty slot (2K)
};

// Counts the document's topics (count_topics, src/cu/topics.cuh:142-158) and builds the key
// table sized for its n0 distinct topics; the map grows (rebuilds) when a new topic would exceed
// the load limit.  The reference starts from the K_d hint (hashmap.cuh:98).
__device__ __forceinline__ DocState open_span(const SweepParams& P, u64 b, u32 L, u32 gdoc, u32* smem) {
  DocState ds;
  const DocLds m = doc_lds(smem, P.K);
  ds.b = uniform_u64(b);
  ds.L = uniform_u32(L);
  ds.cap_max = map_capacity(ds.L, P.K);
  ds.kmax = ds.L < P.K ? ds.L : P.K;
  ds.gdoc = uniform_u32(gdoc);
  ds.hm = map_home_mult(P.K);
  ds.e2 = 2u * P.K;
  ds.occupied = doc_count_topics(m.C, P.K, P.z, ds.b, ds.L);
  ds.cap = map_capacity(ds.occupied, P.K);
  keymap_build(m.Kt, ds.cap, ds.hm, m.C, P.K);
  return ds;
}

__device__ __forceinline__ DocState open_doc(const SweepParams& P, u32 doc, u32* smem) {
  const u64 b = P.doc_off[doc];
  return open_span(P, b, (u32)(P.doc_off[doc + 1] - b), P.doc_base + doc, smem);
}

// Outcome of the draw for one token: the table value of the sparse pick (kNoSlot for a dense draw)
// and its slot.
struct Draw {
  u32 key2;
  u32 slot;
};

// Lane l* of the sparse pick and its exclusive prefix E.
__device__ __forceinline__ u32 pick_lane_only(f32 A, f32 P, f32 tgt) {
  const u64 pos = ballot(A > 0.0f);
  const u64 c = pos & ballot(P > tgt);
  return c ? first_lane(c) : 63u - (u32)__clzll((long long)pos);
}
__device__ __forceinline__ u32 pick_lane(f32 A, f32 P, f32 tgt, f32& E) {
  const u32 ls = pick_lane_only(A, P, tgt);
  E = ls ? readlane_f32(P, ls - 1) : 0.0f;
  return ls;
}

// In-column search over lines read transposed (lane k holds line c0 + k of column ls), one scan
// per 64 lines.  Returns the picked line.
__device__ __forceinline__ u32 column_pick(const u16* Kt, const u16* C, u32 lines, u32 ls, const PhiRow& prow,
                                           f32 E, f32 tgt, u32 e2) {
  const u32 lane = lane_id();
  f32 carry = 0.0f;
  u32 last = 0;
  for (u32 c0 = 0; c0 < lines; c0 += 64) {
    const u32 j = c0 + lane;
    const f32 m = slot_mass(C, j < lines ? (u32)Kt[j * 64 + ls] : e2, prow);
    const f32 Q = wave_scan_f32(m);
    const u64 pos = ballot(m > 0.0f);
    const u64 hit = pos & ballot(E + (carry + Q) > tgt);
    if (hit) return c0 + first_lane(hit);
    if (pos) last = c0 + 63u - (u32)__clzll((long long)pos);
    carry = carry + readlane_f32(Q, 63);
  }
  return last;
}

// A token's values as the draw reads them: lane t of sa, ux and uw holds its sigma_a and its
// sparse/dense and target uniforms (a sweep passes its chunk's per-lane values, so nothing is moved
// to scalar registers; a caller with one token passes the same values in every lane and t = 0).
struct TokenU {
  f32 sa, ux, uw;
  u32 t;
};
// Sparse when u_x * (sigma_a + sigma_b) > sigma_a, evaluated in lane t and read from the lane mask.
__device__ __forceinline__ bool draw_sparse(const TokenU& u, f32 sb) {
  return (ballot(u.ux * (u.sa + sb) > u.sa) >> u.t) & 1u;
}
// The sparse target u_w * sigma_b, evaluated in lane t.
__device__ __forceinline__ f32 draw_target(const TokenU& u, f32 sb) { return readlane_f32(u.uw * sb, u.t); }

// Sparse/dense pick with the NL lines of the table held in registers (one variant per line count;
// e[j] is line j's value in this lane's column): all phi gathers of the token are issued before
// any arithmetic on them.
template <int NL>
__device__ __forceinline__ Draw draw_topic_regs(const u32 (&e)[NL], const u16* C, const PhiRow& prow,
                                                const TokenU& u) {
  // every count read and phi gather of the token issued before any arithmetic on them
  u32 c[NL];
#pragma unroll
  for (int j = 0; j < NL; ++j) c[j] = count_at(C, e[j]);
  f32 f[NL];
  row_values<NL>(prow, e, f);
  f32 m[NL];
#pragma unroll
  for (int j = 0; j < NL; ++j) m[j] = (f32)c[j] * f[j];
  // G_g starts at line g's mass (0 + m == m, masses are never -0); with fewer than four lines the
  // empty groups add +0, which leaves the sum's bits unchanged, so those adds are dropped
  f32 G[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) G[g] = g < NL ? m[g] : 0.0f;
#pragma unroll
  for (int j = 4; j < NL; ++j) G[j & 3] = G[j & 3] + m[j];
  f32 A = G[0];
  if (NL >= 2) A = A + G[1];
  if (NL >= 3) A = A + G[2];
  if (NL >= 4) A = A + G[3];
  const f32 P = wave_scan_f32(A);
  const f32 sb = readlane_f32(P, 63);
  Draw d;
  d.key2 = kNoSlot;
  d.slot = kNoSlot;
  if (!draw_sparse(u, sb)) return d;
  const f32 tgt = draw_target(u, sb);
  const u32 ls = pick_lane_only(A, P, tgt);
  // lane l holds P[l - 1] (0 in lane 0), so lane ls holds E, the exclusive prefix of its column
  const f32 E = wave_shr1_zero(P);
  // every lane scans its own column in registers with the canonical network (NL <= 16 lines lie
  // in the network's first row: offsets 1, 2, 4, 8, entries below the offset kept); lane ls's
  // result is the one used
  f32 q[NL];
#pragma unroll
  for (int j = 0; j < NL; ++j) q[j] = m[j];
#pragma unroll
  for (int off = 1; off < 16; off <<= 1)
#pragma unroll
    for (int j = NL - 1; j >= off; --j) q[j] = q[j] + q[j - off];
  // the picked line and its table value, packed as line << 16 | key2 (key2 < 2^16), so that one
  // readlane of lane ls gives both
  u32 pick = kNoSlot;
#pragma unroll
  for (int j = NL - 1; j >= 0; --j) {
    if (m[j] > 0.0f && E + q[j] > tgt) pick = ((u32)j << 16) | e[j];  // carry 0: E + (0 + q) == E + q
  }
  u32 pk = readlane_u32(pick, ls);
  if (pk == kNoSlot) {  // no line passes the target (rounding): the last line with mass
    u32 last = e[0];
#pragma unroll
    for (int j = 1; j < NL; ++j)
      if (m[j] > 0.0f) last = ((u32)j << 16) | e[j];
    pk = readlane_u32(last, ls);
  }
  d.key2 = pk & 0xFFFFu;
  d.slot = (pk >> 16) * 64 + ls;
  return d;
}

// The register variant with its lines loaded from the table.
template <int NL>
__device__ __forceinline__ Draw draw_topic_regs_lds(const DocLds& m, const PhiRow& prow, const TokenU& u) {
  u32 e[NL];
#pragma unroll
  for (int j = 0; j < NL; ++j) e[j] = m.Kt[j * 64 + lane_id()];
  return draw_topic_regs<NL>(e, m.C, prow, u);
}

// Sparse/dense pick for any number of lines: one pass for the column sums, then for a sparse draw
// the in-column search over column l* read transposed.
__device__ __forceinline__ Draw draw_topic_loop(const u16* Kt, const u16* C, u32 lines, const PhiRow& prow,
                                                const TokenU& u, u32 e2) {
  const u32 lane = lane_id();
  f32 G0 = 0.0f, G1 = 0.0f, G2 = 0.0f, G3 = 0.0f;
  // eight lines per step, so eight gathers are in flight at once; full steps read without a
  // bounds test, the last partial step reads an empty slot past the map's end
  u32 j = 0;
  for (; j + 8 <= lines; j += 8) {
    u32 e[8];
#pragma unroll
    for (u32 k = 0; k < 8; ++k) e[k] = Kt[(j + k) * 64 + lane];
    u32 c[8];
#pragma unroll
    for (u32 k = 0; k < 8; ++k) c[k] = count_at(C, e[k]);
    f32 f[8];
    row_values<8>(prow, e, f);
    f32 m[8];
#pragma unroll
    for (u32 k = 0; k < 8; ++k) m[k] = (f32)c[k] * f[k];
    G0 = G0 + m[0];
    G1 = G1 + m[1];
    G2 = G2 + m[2];
    G3 = G3 + m[3];
    G0 = G0 + m[4];
    G1 = G1 + m[5];
    G2 = G2 + m[6];
    G3 = G3 + m[7];
  }
  if (j < lines) {
    u32 e[8];
#pragma unroll
    for (u32 k = 0; k < 8; ++k) e[k] = j + k < lines ? (u32)Kt[(j + k) * 64 + lane] : e2;
    u32 c[8];
#pragma unroll
    for (u32 k = 0; k < 8; ++k) c[k] = count_at(C, e[k]);
    f32 f[8];
    row_values<8>(prow, e, f);
    f32 m[8];
#pragma unroll
    for (u32 k = 0; k < 8; ++k) m[k] = (f32)c[k] * f[k];
    G0 = G0 + m[0];
    G1 = G1 + m[1];
    G2 = G2 + m[2];
    G3 = G3 + m[3];
    G0 = G0 + m[4];
    G1 = G1 + m[5];
    G2 = G2 + m[6];
    G3 = G3 + m[7];
  }
  const f32 A = combine_groups(G0, G1, G2, G3);
  const f32 P = wave_scan_f32(A);
  const f32 sb = readlane_f32(P, 63);
  Draw d;
  d.key2 = kNoSlot;
  d.slot = kNoSlot;
  if (!draw_sparse(u, sb)) return d;
  const f32 tgt = draw_target(u, sb);
  f32 E;
  const u32 ls = pick_lane(A, P, tgt, E);
  d.slot = column_pick(Kt, C, lines, ls, prow, E, tgt, e2) * 64 + ls;
  d.key2 = Kt[d.slot];
  return d;
}

// Count changes at the end of a token: +1 for its new topic (lane 0) and, when the chunk holds a
// next token, -1 for that token's old topic (lane 1), both in one read and one write of the counts
// (nothing when the two topics are equal).  Reports whether the new topic entered (its count was
// 0) and whether the next token's topic left (its count reached 0).
struct CountStep {
  u32 entered;
  u32 left;
};
__device__ __forceinline__ CountStep count_step(u16* C, u32 new2, u32 next2, u32 has_next) {
  const u32 lane = lane_id();
  const u32 n = has_next ? (next2 == new2 ? 0u : 2u) : 1u;
  // lane 0 the new topic, lane 1 the next token's old topic (every other lane the new topic)
  const u32 a = lane == 1 ? next2 : new2;
  u32 old = 0u;
  if (lane < n) {
    old = count_at(C, a);
    count_put(C, a, old + (lane == 0 ? 1u : 0xFFFFu));  // 16-bit: + 0xFFFF is - 1
  }
  // entered: n != 0 and lane 0's old count was 0; left: n == 2 and lane 1's old count was 1 (the
  // tests as lane masks, bits 0 and 1 taken on the scalar unit)
  const u32 z0 = (u32)ballot(old == 0u), z1 = (u32)ballot(old == 1u);
  CountStep cs;
  cs.entered = z0 & (n | (n >> 1)) & 1u;  // n is 0, 1 or 2: (n | n >> 1) & 1 is n != 0
  cs.left = (z1 >> 1) & (n >> 1) & 1u;    // n >> 1 is n == 2
  return cs;
}

// A token's key-set changes, in the order of the definition: the new topic enters the table (the
// table growing first when it would pass its load limit, rebuilt from the counts, which already
// hold both changes: the layout is a function of the key set), then the next token's old topic
// leaves it.  Returns true when the capacity changed.
__device__ __forceinline__ bool apply_key_changes(const SweepParams& P, const DocLds& m, DocState& ds,
                                                  const CountStep& cs, u32 new2, u32 next2) {
  if (cs.entered) {
    if (map_capacity(ds.occupied + 1, P.K) > ds.cap) {
      const u32 want = 2 * (ds.occupied + 1) < ds.kmax ? 2 * (ds.occupied + 1) : ds.kmax;
      ds.cap = map_capacity(want, P.K);
      ds.occupied = ds.occupied + 1 - cs.left;
      wave_sync();
      keymap_build(m.Kt, ds.cap, ds.hm, m.C, P.K);
      return true;
    }
    keymap_insert(m.Kt, ds.cap, ds.hm, ds.e2, new2);
    ++ds.occupied;
  }
  if (cs.left) {
    keymap_erase(m.Kt, ds.cap, ds.hm, ds.e2, next2);
    --ds.occupied;
  }
  return false;
}

// Removal of the first token of a chunk (the others are removed at the end of the token before).
__device__ __forceinline__ void remove_first(const DocLds& m, DocState& ds, u32 z2) {
  u32 old = 0;
  wave_sync();
  if (lane_id() == 0) {
    old = count_at(m.C, z2);
    count_put(m.C, z2, old - 1u);
  }
  if (readlane_u32(old, 0) == 1u) {
    keymap_erase(m.Kt, ds.cap, ds.hm, ds.e2, z2);
    --ds.occupied;
  }
}

// Per-lane values of the 64 tokens of one chunk: lane l holds token c0 + l.
struct Chunk {
  u32 ww;      // word
  u32 zz2;     // 2 * topic before the pass
  u32 slot;    // alias slot of a dense draw (from r.y); kCountsFitBit: the word's count row fits
  f32 uz;      // uniform of the dense draw's keep test (r.z)
  f32 sa;      // sigma_a of the word
  f32 ux;      // uniform of the sparse/dense choice
  f32 uw;      // uniform of the sparse target
  u32 znew;    // topic after the pass (lanes below the current token)
};

constexpr u32 kCountsFitBit = 0x80000000u;

// Tokens t .. nt-1 of a chunk while the table has NL lines (NL = 1..8: lines held in registers;
// 0: read per token).  Returns the next token: nt, or the token after the one whose insertion
// grew the map.  All variants evaluate the same operations in the same order: a variant for fewer
// lines drops only lines that would hold no mass.
template <int NL, bool CR>
__device__ __forceinline__ u32 sweep_tokens(const SweepParams& P, const DocLds& m, DocState& ds, Chunk& ch, u32 t,
                                            u32 nt, u32& visited) {
  const u32 lane = lane_id();
  const u32 K = P.K;
  u32 e[NL > 0 ? NL : 1];
  if constexpr (NL > 0) {
    wave_sync();
#pragma unroll
    for (int j = 0; j < NL; ++j) e[j] = m.Kt[j * 64 + lane];
  }
  u32 cur2 = readlane_u32(ch.zz2, t);  // token t's topic before the pass (carried from next2 below)
  // per chunk, as lane masks: words with mass, and words whose count rows the sweep reads
  const u64 live = ballot(ch.sa > 0.0f);
  const u64 fit = CR ? ballot((ch.slot & kCountsFitBit) != 0u) : 0ull;
  do {  // entered with t < nt
    const u32 wt = readlane_u32(ch.ww, t);
    u32 new2 = cur2;  // a word with an all-zero phi row keeps its topic
    wave_sync();  // the previous token's count writes before this token's reads
    if ((live >> t) & 1u) {
      visited += ds.occupied;
      const PhiRow prow = CR && ((fit >> t) & 1u) ? phi_count_row(P.phi_counts + (u64)wt * K, K, m.recip)
                                                  : phi_row(P.phi + (u64)wt * K, K);
      const TokenU u{ch.sa, ch.ux, ch.uw, t};
      Draw d;
      if constexpr (NL > 0) d = draw_topic_regs<NL>(e, m.C, prow, u);
      else d = draw_topic_loop(m.Kt, m.C, uniform_u32(ds.cap / kMapLine), prow, u, ds.e2);
      if (d.key2 != kNoSlot) {
        new2 = d.key2;
      } else {  // dense: the word's alias slot is read only now (about one token in eight)
        const u32 sl = readlane_u32(ch.slot, t) & ~kCountsFitBit;
        const AliasSlot a = P.alias[(u64)wt * K + sl];
        new2 = uniform_u32(2u * (readlane_f32(ch.uz, t) < a.prob ? sl : a.alias));  // (the load is uniform)
      }
    }
    if (lane == t) ch.znew = new2 >> 1;
    const u32 has_next = t + 1 < nt ? 1u : 0u;
    const u32 next2 = readlane_u32(ch.zz2, t + 1 < 64 ? t + 1 : 63);
    cur2 = next2;
    const CountStep cs = count_step(m.C, new2, next2, has_next);
    if (cs.entered | cs.left) {
      wave_sync();
      if (apply_key_changes(P, m, ds, cs, new2, next2)) return t + 1;
      if constexpr (NL > 0) {
#pragma unroll
        for (int j = 0; j < NL; ++j) e[j] = m.Kt[j * 64 + lane];
      }
    }
  } while (++t < nt);
  return nt;
}

// Runs tokens t .. nt-1 of a chunk with the variant for the table's current line count (register
// variants up to 8 lines, within the kernel's LMAX, which caps its registers; larger tables use
// the line loop).  The choice is made again only when an insertion grows the table.
template <int LMAX, bool CR>
__device__ __forceinline__ u32 sweep_tokens_any(const SweepParams& P, const DocLds& m, DocState& ds, Chunk& ch,
                                                u32 t, u32 nt, u32& visited) {
  switch (ds.cap / kMapLine) {  // a map always has at least one line
    case 1: if (LMAX >= 1) return sweep_tokens<1, CR>(P, m, ds, ch, t, nt, visited); break;
    case 2: if (LMAX >= 2) return sweep_tokens<2, CR>(P, m, ds, ch, t, nt, visited); break;
    case 3: if (LMAX >= 4) return sweep_tokens<3, CR>(P, m, ds, ch, t, nt, visited); break;
    case 4: if (LMAX >= 4) return sweep_tokens<4, CR>(P, m, ds, ch, t, nt, visited); break;
    case 5: if (LMAX >= 8) return sweep_tokens<5, CR>(P, m, ds, ch, t, nt, visited); break;
    case 6: if (LMAX >= 8) return sweep_tokens<6, CR>(P, m, ds, ch, t, nt, visited); break;
    case 7: if (LMAX >= 8) return sweep_tokens<7, CR>(P, m, ds, ch, t, nt, visited); break;
    case 8: if (LMAX >= 8) return sweep_tokens<8, CR>(P, m, ds, ch, t, nt, visited); break;
    default: break;
  }
  return sweep_tokens<0, CR>(P, m, ds, ch, t, nt, visited);
}

// One pass over a document's tokens with Philox counters (token, gdoc, iter, tag).  visited gains
// the sum over sampled tokens of the map's distinct topics (at most 65535 * 65535 per pass, so
// the pass counts in 32 bits).
template <int LMAX, bool CR = false>
__device__ __forceinline__ void sweep_pass(const SweepParams& P, const DocLds& m, DocState& ds, u32 iter, u32 tag,
                                           u64& visited) {
  const u32 lane = lane_id();
  const u32 K = P.K;
  u32 pass_visited = 0;
  for (u32 c0 = 0; c0 < ds.L; c0 += 64) {
    const u32 i = c0 + lane;
    const bool valid = i < ds.L;
    Chunk ch;
    ch.ww = valid ? P.w[ds.b + i] : 0;
    const u32 zz = valid ? (u32)P.z[ds.b + i] : 0;
    ch.zz2 = 2u * zz;
    const u32x4 r = philox_tagged(i, ds.gdoc, iter, tag, 0, P.k0, P.k1);
    // per-token values that do not depend on the map: computed for the whole chunk
    ch.sa = valid ? P.sigma_a[ch.ww] : 0.0f;
    ch.slot = bounded(r.y, K);
    if (CR && valid && P.counts_fit[ch.ww]) ch.slot |= kCountsFitBit;
    ch.uz = u24(r.z);
    ch.ux = u24(r.x);
    ch.uw = u24(r.w);
    ch.znew = zz;
    const u32 nt = ds.L - c0 < 64 ? ds.L - c0 : 64;
    remove_first(m, ds, readlane_u32(ch.zz2, 0));
    for (u32 t = 0; t < nt;) t = sweep_tokens_any<LMAX, CR>(P, m, ds, ch, t, nt, pass_visited);
    if (valid) P.z[ds.b + i] = (u16)ch.znew;
  }
  visited += pass_visited;
}

// Fold-in carries the map from sweep to sweep.  Once the document's topics have thinned out, the
// key table is rebuilt at the capacity its keys need (the layout is a function of the key set and
// the capacity), so the later sweeps scan fewer lines.
__device__ __forceinline__ void shrink_doc_map(const SweepParams& P, const DocLds& m, DocState& ds) {
  const u32 want = map_capacity(ds.occupied, P.K);
  if (want < ds.cap) {
    ds.cap = want;
    wave_sync();
    keymap_build(m.Kt, ds.cap, ds.hm, m.C, P.K);
  }
}

// All sweeps of a fold-in (counters (token, request, s, FOLD_SWEEP), s = 1..sweeps).
__device__ __forceinline__ void foldin_sweeps(const SweepParams& P, u32* smem, DocState& ds, u32 sweeps,
                                              u64& visited) {
  const DocLds m = doc_lds(smem, P.K);
  for (u32 sweep = 1; sweep <= sweeps; ++sweep) {
    if (sweep > 1) shrink_doc_map(P, m, ds);
    sweep_pass<8>(P, m, ds, sweep, kTagFoldSweep, visited);
  }
}

__device__ __forceinline__ u32 sweep_iteration(const SweepParams& P) { return P.iter_ptr ? *P.iter_ptr : P.iter; }

template <int LMAX, bool CR = false>
__device__ __forceinline__ void sweep_doc(const SweepParams& P, u32 doc, u32* smem,
                                          const float2* recip = nullptr) {
  DocState ds = open_doc(P, doc, smem);
  u64 visited = 0;
  sweep_pass<LMAX, CR>(P, doc_lds(smem, P.K, recip), ds, sweep_iteration(P), kTagSweep, visited);
  if (lane_id() == 0 && P.visited) atomicAdd(P.visited, (unsigned long long)visited);
}

// One wavefront per document, documents taken from P.doc_order (a length bucket) or in order.
template <int LMAX>
__global__ __launch_bounds__(64) void k_sweep_lines(SweepParams P) {
  extern __shared__ u32 sweep_smem[];
  const u32 doc = P.doc_order ? P.doc_order[blockIdx.x] : blockIdx.x;
  if (doc >= P.n_docs) return;
  sweep_doc<LMAX>(P, doc, sweep_smem);
}

// LMAX = 1, 2, 4, 8: the register variants up to that many lines; 0: any map (register variants up
// to 8 lines are then not compiled in, so larger maps keep the registers of the line-loop form).
template __global__ void k_sweep_lines<0>(SweepParams);
template __global__ void k_sweep_lines<1>(SweepParams);
template __global__ void k_sweep_lines<2>(SweepParams);
template __global__ void k_sweep_lines<4>(SweepParams);
template __global__ void k_sweep_lines<8>(SweepParams);

// Persistent form: a fixed grid of wavefronts takes documents from a queue (P.doc_order, longest
// first) through an atomic head.  Every wavefront leaves once the head passes n_items, so the
// grid always drains.  Results are identical to the per-bucket launches: the randomness of a
// document depends on its global id only.
// Registers are budgeted for eight wavefronts per SIMD (the compiler would otherwise stop at seven
// for its scalar registers, and spilling some of them to vector lanes costs less than the lost
// wavefront: the sweep waits on its gathers and LDS more than it issues).
template <int LMAX>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) void k_sweep_persistent(SweepParams P,
                                                                                                u32* head,
                                                                                                u32 n_items) {
  extern __shared__ u32 sweep_smem[];
  for (;;) {
    u32 item = 0;
    if (lane_id() == 0) item = atomicAdd(head, 1u);
    item = readlane_u32(item, 0);
    if (item >= n_items) return;
    sweep_doc<LMAX>(P, P.doc_order[item], sweep_smem);
  }
}
template __global__ void k_sweep_persistent<8>(SweepParams, u32*, u32);

// The persistent sweep reading u16 count rows for the words whose counts fit (half the bytes of a
// row: the sweep's time follows the row lines it misses in L2).  The block's wavefronts share one
// LDS copy of topic_recip, loaded before any of them starts; after that the wavefronts run
// independently, as in k_sweep_persistent.
__global__ __launch_bounds__(64 * kCountsBlockWaves) __attribute__((amdgpu_waves_per_eu(8))) void
k_sweep_persistent_counts(SweepParams P, u32* head, u32 n_items, u32 wave_lds) {
  extern __shared__ u32 sweep_smem[];
  float2* recip = reinterpret_cast<float2*>(sweep_smem);
  for (u32 k = threadIdx.x; k <= P.K; k += blockDim.x) recip[k] = P.topic_recip[k];
  __syncthreads();
  const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: the state's address in SGPRs
  u32* smem = sweep_smem + (topic_recip_lds_bytes(P.K) + (u64)wave * wave_lds) / 4;
  for (;;) {
    u32 item = 0;
    if (lane_id() == 0) item = atomicAdd(head, 1u);
    item = readlane_u32(item, 0);
    if (item >= n_items) return;
    sweep_doc<8, true>(P, P.doc_order[item], smem, recip);
  }
}

__global__ void k_topic_recip(const unsigned long long* __restrict__ topic_sum, float2* __restrict__ out, u32 K) {
  const u32 k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > K) return;
  float2 v = make_float2(0.0f, 1.0f);
  if (k < K) {
    const f32 d = (f32)(f64)topic_sum[k];  // the divisor the f32 rows were normalised by
    if (d > 0.0f) v = make_float2(1.0f / d, d);
  }
  out[k] = v;
}

__global__ void k_check_counts(const f32* __restrict__ phi, const u16* __restrict__ counts,
                               const u8* __restrict__ fit, const float2* __restrict__ recip, u32 V, u32 K,
                               unsigned long long* bad) {
  const u64 n = (u64)V * K;
  u32 mine = 0;
  for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
    const u32 v = (u32)(i / K), k = (u32)(i - (u64)v * K);
    if (!fit[v]) continue;
    const f32 a = phi_from_count(counts[i], recip[k].x, recip[k].y);
    mine += __builtin_bit_cast(u32, a) != __builtin_bit_cast(u32, phi[i]) ? 1u : 0u;
  }
  if (mine) atomicAdd(bad, (unsigned long long)mine);
}

// Persistent form of the line-loop sweep for long documents (maps past the register variants),
// without the eight-wavefront register budget: the line loop then keeps its scalar state in
// scalar registers (spilling it to vector lanes lengthens each token's dependent chain, which is
// what a long document's sweep time is made of).
__global__ __launch_bounds__(64) void k_sweep_persistent_long(SweepParams P, u32* head, u32 n_items) {
  extern __shared__ u32 sweep_smem[];
  for (;;) {
    u32 item = 0;
    if (lane_id() == 0) item = atomicAdd(head, 1u);
    item = readlane_u32(item, 0);
    if (item >= n_items) return;
    sweep_doc<0>(P, P.doc_order[item], sweep_smem);
  }
}

// ---------------------------------------------------------------------------------------------
// Block-per-document sweep for long documents: four wavefronts share one map.  Wave 0 owns the
// counts and the key table (removal, addition, growth) and the in-column search; wave g sums line
// group g (lines j = g mod 4), which is exactly the grouping of the canonical order, so the result
// is bit-identical to the wavefront-per-document sweep.
// ---------------------------------------------------------------------------------------------
constexpr u32 kBlockWaves = 4;

__global__ __launch_bounds__(256) void k_sweep_block(SweepParams P, u32* head, u32 n_items) {
  extern __shared__ u32 sweep_smem[];
  __shared__ f32 grp[kBlockWaves][64];
  __shared__ u32 sh_cap;
  __shared__ u32 sh_item;
  const u32 lane = lane_id();
  const u32 wave = threadIdx.x >> 6;
  const u32 K = P.K;
  const DocLds m = doc_lds(sweep_smem, K);
  const u32 e2 = 2u * K;
  for (;;) {
    if (threadIdx.x == 0) sh_item = head ? atomicAdd(head, 1u) : (blockIdx.x < n_items ? blockIdx.x : n_items);
    __syncthreads();
    const u32 item = sh_item;
    __syncthreads();
    if (item >= n_items) return;
    const u32 doc = P.doc_order ? P.doc_order[item] : item;
    DocState ds;
    if (wave == 0) ds = open_doc(P, doc, sweep_smem);
    const u64 b = P.doc_off[doc];
    const u32 L = (u32)(P.doc_off[doc + 1] - b);
    const u32 gdoc = P.doc_base + doc;
    const u32 iter = sweep_iteration(P);
    u64 visited = 0;
    for (u32 c0 = 0; c0 < L; c0 += 64) {
      const u32 i = c0 + lane;
      const bool valid = i < L;
      const u32 ww = valid ? P.w[b + i] : 0;
      const u32 zz = valid ? (u32)P.z[b + i] : 0;
      const u32x4 r = philox_tagged(i, gdoc, iter, kTagSweep, 0, P.k0, P.k1);
      const f32 sa_l = valid ? P.sigma_a[ww] : 0.0f;
      u32 dense_l = 0;
      if (wave == 0) {
        const u32 slot_l = bounded(r.y, K);
        AliasSlot al_l;
        al_l.prob = 0.0f;
        al_l.alias = 0;
        if (valid) al_l = P.alias[(u64)ww * K + slot_l];
        dense_l = u24(r.z) < al_l.prob ? slot_l : al_l.alias;
        remove_first(m, ds, 2u * readlane_u32(zz, 0));
        if (lane == 0) sh_cap = ds.cap;
      }
      __syncthreads();
      u32 znew_lane = zz;
      const u32 nt = L - c0 < 64 ? L - c0 : 64;
      for (u32 t = 0; t < nt; ++t) {
        const u32 wt = readlane_u32(ww, t);
        const f32 sa = readlane_f32(sa_l, t);
        u32 new2 = 2u * readlane_u32(zz, t);
        if (sa > 0.0f) {
          const u32 lines = sh_cap / kMapLine;
          const PhiRow prow = phi_row(P.phi + (u64)wt * K, K);
          // wave g sums line group g of every column (four of its lines per step, gathers in
          // flight together, added in line order)
          f32 G = 0.0f;
          for (u32 j = wave; j < lines; j += 4 * kBlockWaves) {
            u32 e[4];
#pragma unroll
            for (u32 k = 0; k < 4; ++k) {
              const u32 jj = j + k * kBlockWaves;
              e[k] = jj < lines ? (u32)m.Kt[jj * 64 + lane] : e2;
            }
            f32 mm[4];
#pragma unroll
            for (u32 k = 0; k < 4; ++k) mm[k] = slot_mass(m.C, e[k], prow);
#pragma unroll
            for (u32 k = 0; k < 4; ++k) G = G + mm[k];
          }
          grp[wave][lane] = G;
          __syncthreads();
          if (wave == 0) {
            const f32 A = combine_groups(grp[0][lane], grp[1][lane], grp[2][lane], grp[3][lane]);
            const f32 Ps = wave_scan_f32(A);
            const f32 sb = readlane_f32(Ps, 63);
            visited += ds.occupied;
            if (u24(readlane_u32(r.x, t)) * (sa + sb) > sa) {
              const f32 tgt = u24(readlane_u32(r.w, t)) * sb;
              f32 E;
              const u32 ls = pick_lane(A, Ps, tgt, E);
              new2 = m.Kt[column_pick(m.Kt, m.C, lines, ls, prow, E, tgt, e2) * 64 + ls];
            } else {
              new2 = 2u * readlane_u32(dense_l, t);
            }
          }
        }
        if (wave == 0) {
          new2 = uniform_u32(new2);
          if (lane == t) znew_lane = new2 >> 1;
          const u32 has_next = t + 1 < nt ? 1u : 0u;
          const u32 next2 = 2u * readlane_u32(zz, t + 1 < 64 ? t + 1 : 63);
          const CountStep cs = count_step(m.C, new2, next2, has_next);
          if (cs.entered | cs.left) {
            wave_sync();
            apply_key_changes(P, m, ds, cs, new2, next2);
          }
          if (lane == 0) sh_cap = ds.cap;
        }
        __syncthreads();
      }
      if (wave == 0 && valid) P.z[b + i] = (u16)znew_lane;
      __syncthreads();
    }
    if (wave == 0 && lane == 0 && P.visited) atomicAdd(P.visited, (unsigned long long)visited);
    if (!head) return;
  }
}

// ---------------------------------------------------------------------------------------------
// Fold-in: topics of a new document against a fixed model.  Each request (a span of the token
// arena) starts from its words' alias tables -- a draw from p(k | w) proportional to phi[w][k],
// uniform for an all-zero row -- with counters (token, request id, 0, FOLD_INIT), then runs
// `sweeps` passes with counters (token, request id, s, FOLD_SWEEP) and the map carried from pass
// to pass.  A request's result depends only on its words and its id, never on its batch.
// When done flags are given, the final topics go to the host-visible output and then the
// request's flag is raised (system-scope fence in between).
// ---------------------------------------------------------------------------------------------
template <int LMAX>
__global__ __launch_bounds__(64) void k_foldin(SweepParams P, FoldParams F) {
  extern __shared__ u32 fold_smem[];
  const u32 lane = lane_id();
  const u32 slot = F.slots ? F.slots[blockIdx.x] : blockIdx.x;
  if (slot >= F.n_slots) return;
  const u64 b = F.span_off[slot];
  const u32 L = F.span_len[slot];
  const u32 rid = F.req_id[slot];
  const u32 K = P.K;
  for (u32 i = lane; i < L; i += 64) {
    const u32 w = P.w[b + i];
    const u32x4 r = philox_tagged(i, rid, 0, kTagFoldInit, 0, P.k0, P.k1);
    const u32 z = P.sigma_a[w] > 0.0f ? draw_alias_slot(P.alias + (u64)w * K, K, r.y, r.z) : bounded(r.x, K);
    P.z[b + i] = (u16)z;
  }
  __syncthreads();
  DocState ds = open_span(P, b, L, rid, fold_smem);
  u64 visited = 0;
  foldin_sweeps(P, fold_smem, ds, F.sweeps, visited);
  if (F.z_out) {
    __syncthreads();
    for (u32 i = lane; i < L; i += 64) F.z_out[b + i] = P.z[b + i];
  }
  if (F.done) {
    __threadfence_system();
    if (lane == 0) {
      __atomic_store_n(&F.done[slot], F.done_value, __ATOMIC_RELEASE);
    }
  }
  if (lane == 0 && P.visited) atomicAdd(P.visited, (unsigned long long)visited);
}
template __global__ void k_foldin<8>(SweepParams, FoldParams);

// ---------------------------------------------------------------------------------------------
// Serving (continuous batching).
//   k_serve_feed   one launch per flush, after the host's copies of the flush's words (into the
//                  arena) and records on the same stream: writes each record into its slot's line,
//                  then publishes ring entry base + i (slot + 1) with an agent-scope release and
//                  raises the announced count `tail`;
//   k_serve_worker a resident pool: each wavefront claims the next announced entry with a
//                  compare-and-swap on `head`, takes the entry once published, checks the record
//                  (it names the claimed entry) and the words (their sum is the recorded one, all
//                  ids below V), runs the request's fold-in (the steps of k_foldin), writes the
//                  topics and their checksum to host memory and raises the completion flag.
// Control flow in the worker is uniform: every loop and branch tests a scalar, and lane 0 alone
// only ever issues a single atomic, so no part of the wave can be left behind in a loop.  A
// worker stops claiming once the pool's lifetime has passed or `shutdown` is raised and waits for
// an entry, a record or words at most that long, so every wave reaches its exit.  Data that does
// not validate in time is never used: the request's flag reports the failure.
// ---------------------------------------------------------------------------------------------
// Polling read of a shared counter: coherent (agent scope) but relaxed, so an idle wave does not
// invalidate its compute unit's caches under the waves that are working there; the one acquire
// fence follows a successful claim.
__device__ __forceinline__ u32 scalar_load(const u32* p) {
  return uniform_u32(__hip_atomic_load(const_cast<u32*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}
// A load that bypasses the vector and L2 caches (system scope), for data another agent wrote.
__device__ __forceinline__ u32 fresh_load(const u32* p) {
  return __hip_atomic_load(const_cast<u32*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ __launch_bounds__(64) void k_serve_feed(ServeFeed Fd, ServeRing R) {
  const u32 k = blockIdx.x * 64 + threadIdx.x;
  if (k >= Fd.count) return;
  const u32* m = Fd.meta + (u64)k * kServeMetaWords;
  const u32 slot = fresh_load(&m[0]);
  if (slot >= R.ring_size) return;  // never from the host; the request would then time out
  const u32 idx = Fd.base + k;
  u32* rec = R.rec + (u64)slot * kServeRecWords;
  for (u32 q = 1; q < kServeMetaWords; ++q) rec[q] = fresh_load(&m[q]);
  __threadfence();
  rec[0] = idx + 1u;  // the record is complete once it names this entry
  __threadfence();    // record before the entry
  __hip_atomic_exchange(&R.ring[idx % R.ring_size], slot + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_fetch_max(R.tail, idx + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(64) void k_serve_worker(SweepParams P, FoldParams F, ServeRing R) {
  extern __shared__ u32 serve_smem[];
  const u32 lane = lane_id();
  const u32 K = P.K;
  const u64 start = wall_clock64();
  for (;;) {
    // claim the next announced entry
    u32 item = kNoSlot;
    for (u32 nap = 1;;) {
      if (scalar_load(R.shutdown) != 0u || wall_clock64() - start > R.lifetime_ticks) break;
      const u32 t = scalar_load(R.tail);
      const u32 h = scalar_load(R.head);
      // head is read after tail, so other waves may have moved it past this tail: only an entry
      // below the tail read here is announced (signed difference, the counters wrap)
      if ((int)(t - h) > 0) {
        u32 won = 0;
        if (lane == 0) {
          u32 e = h;
          won = __hip_atomic_compare_exchange_strong(R.head, &e, h + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)
                    ? 1u
                    : 0u;
        }
        if (readlane_u32(won, 0)) {
          item = h;
          break;
        }
        continue;  // another wave took it: retry at once
      }
      for (u32 q = 0; q < nap; ++q) __builtin_amdgcn_s_sleep(16);
      nap = nap < 32 ? 2 * nap : 32;
    }
    if (item == kNoSlot) return;
    // take the entry (announced before its feeder has published it); taking it clears it for the
    // next lap of the ring
    u32 slot = kNoSlot;
    for (const u64 t0 = wall_clock64();;) {
      u32 v = 0;
      if (lane == 0) v = __hip_atomic_exchange(&R.ring[item % R.ring_size], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      v = readlane_u32(v, 0);
      if (v != 0u) {
        slot = v - 1u;
        break;
      }
      if (scalar_load(R.shutdown) != 0u || wall_clock64() - t0 > R.lifetime_ticks) break;
      __builtin_amdgcn_s_sleep(8);
    }
    if (slot >= R.ring_size) return;  // no entry in time: the host reports the request as lost
    __atomic_thread_fence(__ATOMIC_ACQUIRE);  // the entry was published with a release
    // the record, until it names the claimed entry
    const u32* rec = R.rec + (u64)slot * kServeRecWords;
    u32 rv = 0;
    bool ok = false;
    for (const u64 t0 = wall_clock64();;) {
      rv = lane < kServeMetaWords ? fresh_load(&rec[lane]) : 0u;
      ok = readlane_u32(rv, 0) == item + 1u;
      if (ok || wall_clock64() - t0 > R.lifetime_ticks) break;
      __builtin_amdgcn_s_sleep(4);
    }
    const u32 L = readlane_u32(rv, 1);
    const u32 rid = readlane_u32(rv, 2);
    const u64 b = (u64)readlane_u32(rv, 3) | ((u64)readlane_u32(rv, 4) << 32);
    const u32 want_sum = readlane_u32(rv, 5);
    ok = ok && L >= 1 && L <= kMaxDocLen && b + L <= R.arena;
    // the words, until they sum to the recorded value (and all ids lie below V)
    u32 fp = 0;
    if (ok) {
      ok = false;
      for (const u64 t0 = wall_clock64();;) {
        fp = 0;
        u32 out_of_range = 0;
        for (u32 i = lane; i < L; i += 64) {
          const u32 w = fresh_load(&P.w[b + i]);
          fp += w * (2u * i + 1u);
          out_of_range |= w >= R.V ? 1u : 0u;
        }
        for (u32 off = 32; off > 0; off >>= 1) fp += __shfl_xor(fp, off, 64);
        const bool range_ok = ballot(out_of_range != 0u) == 0ull;
        ok = fp == want_sum && range_ok;
        if (ok || wall_clock64() - t0 > R.lifetime_ticks) break;
        __builtin_amdgcn_s_sleep(4);
      }
    }
    __threadfence();  // the plain reads of the arena below come after the validated ones
    u32 flag = kServeFailed, zsum = 0;
    if (ok) {
      for (u32 i = lane; i < L; i += 64) {
        const u32 w = P.w[b + i];
        const u32x4 r = philox_tagged(i, rid, 0, kTagFoldInit, 0, P.k0, P.k1);
        const u32 z = P.sigma_a[w] > 0.0f ? draw_alias_slot(P.alias + (u64)w * K, K, r.y, r.z) : bounded(r.x, K);
        P.z[b + i] = (u16)z;
      }
      wave_sync();
      DocState ds = open_span(P, b, L, rid, serve_smem);
      u64 visited = 0;
      foldin_sweeps(P, serve_smem, ds, F.sweeps, visited);
      wave_sync();
      // the topics go to host memory with system-scope stores and their checksum in the record
      for (u32 i = lane; i < L; i += 64) {
        const u32 z = P.z[b + i];
        zsum += (z + 1u) * (2u * i + 1u);
        __hip_atomic_store(&F.z_out[b + i], (u16)z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
      }
      for (u32 off = 32; off > 0; off >>= 1) zsum += __shfl_xor(zsum, off, 64);
      flag = serve_fingerprint(fp, L, rid);
    }
    // what this wave read and wrote, then the completion flag (the fingerprint of what it read)
    u32* dr = R.done_rec + (u64)slot * kServeRecWords;
    const u32 out_rec[8] = {item + 1u, L, rid, (u32)b, (u32)(b >> 32), fp, zsum, ok ? 1u : 0u};
    u32 v = 0;
#pragma unroll
    for (u32 q = 0; q < 8; ++q)
      if (lane == q) v = out_rec[q];
    if (lane < 8) __hip_atomic_store(&dr[lane], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __threadfence_system();
    if (lane == 0) __hip_atomic_store(&F.done[slot], flag, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// ---------------------------------------------------------------------------------------------
// n_wk reduction: z to word order, then per-word histograms, every pass reading and writing HBM
// in order.
//   k_tile_reorder  one workgroup per tile of kCountTile consecutive entries of `in`: each entry
//                   goes to LDS at its rank inside the tile, and the tile is written out in rank
//                   order to out[dst[...]], a run at a time.  Ranks and destinations are fixed by
//                   the words, so they are built once per corpus.  It runs twice: z (document
//                   order) -> z1 (word groups: contiguous word ranges of about N / kCountGroups
//                   tokens; tiles write runs of about kCountTile / kCountGroups tokens), then
//                   z1 -> zs (word-sorted: words ascending, each word's tokens in corpus order;
//                   a tile of z1 holds one or two groups, so each word writes about one run per
//                   tile of its group).  One pass straight to word order would write one run per
//                   token of every word rarer than one token per tile.
//   k_count         one wavefront per task (a run of at most kCountTask tokens of one word, a
//                   contiguous range of zs): LDS histogram over K, nonzero bins added to n_wk.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_tile_reorder(const u16* __restrict__ in, TileRuns R, u16* __restrict__ out,
                                                       u64 n) {
  constexpr u32 kWords = kCountTile / 32;
  __shared__ u16 tile[kCountTile];
  __shared__ u32 wbits[kWords];   // run starts of this tile
  __shared__ u32 wpref[kWords];   // runs starting before each word
  __shared__ u32 wtot[1024 / 64];
  const u64 t0 = (u64)blockIdx.x * kCountTile;  // tiles start 64 KiB apart in `in` and `rank`
  const u32 m = (u32)(n - t0 < kCountTile ? n - t0 : kCountTile);
  const u32 m4 = m & ~3u;
  const u16* rank = R.rank;
  // exclusive prefix counts of run starts per bitmap word (one word per thread: a block scan)
  static_assert(kWords == 1024, "one bitmap word per thread");
  const u32 wb = R.bits[(u64)blockIdx.x * kWords + threadIdx.x];
  // four entries per lane with 8-byte loads of values and ranks; a full tile (launched with 1024
  // threads) has all eight of each lane's load pairs in flight before its LDS stores
  auto put4 = [&](uint2 v, uint2 r) {
    tile[r.x & 0xFFFFu] = (u16)(v.x & 0xFFFFu);
    tile[r.x >> 16] = (u16)(v.x >> 16);
    tile[r.y & 0xFFFFu] = (u16)(v.y & 0xFFFFu);
    tile[r.y >> 16] = (u16)(v.y >> 16);
  };
  constexpr u32 kFullPer = kCountTile / (4 * 1024);
  if (m == kCountTile && blockDim.x == 1024) {
    uint2 v[kFullPer], r[kFullPer];
#pragma unroll
    for (u32 i = 0; i < kFullPer; ++i) {
      const u32 l = 4 * (threadIdx.x + i * 1024);
      v[i] = *reinterpret_cast<const uint2*>(in + t0 + l);
      r[i] = *reinterpret_cast<const uint2*>(rank + t0 + l);
    }
#pragma unroll
    for (u32 i = 0; i < kFullPer; ++i) put4(v[i], r[i]);
  } else {
    for (u32 l = 4 * threadIdx.x; l < m4; l += 4 * blockDim.x)
      put4(*reinterpret_cast<const uint2*>(in + t0 + l), *reinterpret_cast<const uint2*>(rank + t0 + l));
    for (u32 l = m4 + threadIdx.x; l < m; l += blockDim.x) tile[rank[t0 + l]] = in[t0 + l];
  }
  wbits[threadIdx.x] = wb;
  u32 c = (u32)__popc(wb);
  const u32 lane = lane_id();
  const u32 wave = threadIdx.x >> 6;
  for (u32 off = 1; off < 64; off <<= 1) {
    const u32 y = (u32)__shfl_up((int)c, off, 64);
    if (lane >= off) c += y;
  }
  if (lane == 63) wtot[wave] = c;
  __syncthreads();
  u32 before = 0;
  for (u32 q = 0; q < wave; ++q) before += wtot[q];
  wpref[threadIdx.x] = before + c - (u32)__popc(wb);
  __syncthreads();
  // consecutive lanes write consecutive ranks (a run's entries go to consecutive places); eight
  // independent places per lane in flight
  const uint2* runs = R.runs + R.run_off[blockIdx.x];
  auto place = [&](u32 j) -> u32 {
    const u32 wj = j >> 5;
    const u32 below = wbits[wj] & (0xFFFFFFFFu >> (31u - (j & 31u)));  // starts at ranks <= j in the word
    const uint2 run = runs[wpref[wj] + (u32)__popc(below) - 1u];          // rank 0 starts a run
    return run.y + (j - run.x);
  };
  const u32 B = blockDim.x;
  u32 j = threadIdx.x;
  for (; j + 7 * B < m; j += 8 * B) {
    u32 d[8];
#pragma unroll
    for (u32 i = 0; i < 8; ++i) d[i] = place(j + i * B);
#pragma unroll
    for (u32 i = 0; i < 8; ++i) out[d[i]] = tile[j + i * B];
  }
  for (; j < m; j += B) out[place(j)] = tile[j];
}

__global__ void k_count(const CountTask* __restrict__ tasks, u32 n_tasks, const u16* __restrict__ zs,
                        u32* __restrict__ n, u32 K, u32 waves_per_block) {
  extern __shared__ u32 count_smem[];
  const u32 lane = lane_id();
  const u32 wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // uniform: the task's load is scalar
  const u32 task = blockIdx.x * waves_per_block + wave;
  if (task >= n_tasks) return;
  const CountTask tk = tasks[task];
  // one histogram of K bins per wavefront, K rounded up to a multiple of 4 so that every
  // histogram starts 16-byte aligned; bins are cleared four at a time
  const u32 Kp = (K + 3u) & ~3u;
  u32* H = count_smem + (u64)wave * Kp;
  for (u32 k = 4 * lane; k < Kp; k += 256) *reinterpret_cast<uint4*>(&H[k]) = make_uint4(0u, 0u, 0u, 0u);
  wave_sync();
  // sixteen loads of 64 topics in flight before their LDS adds (one load and add at a time leaves
  // each step waiting on the load)
  for (u32 p = lane; p < tk.len; p += 16 * 64) {
    u32 zb[16];
#pragma unroll
    for (u32 b = 0; b < 16; ++b) zb[b] = p + 64 * b < tk.len ? (u32)zs[tk.p0 + p + 64 * b] : 0u;
#pragma unroll
    for (u32 b = 0; b < 16; ++b)
      if (p + 64 * b < tk.len) atomicAdd(&H[zb[b]], 1u);
  }
  wave_sync();
  // bins read one per lane, consecutive, so each atomic instruction covers two cache lines; four
  // bins per lane read before their atomics
  u32* row = n + (u64)tk.word * K;
  for (u32 k = lane; k < K; k += 4 * 64) {
    u32 c[4];
#pragma unroll
    for (u32 j = 0; j < 4; ++j) c[j] = k + 64 * j < K ? H[k + 64 * j] : 0u;
#pragma unroll
    for (u32 j = 0; j < 4; ++j)
      if (c[j]) atomicAdd(&row[k + 64 * j], c[j]);
  }
}

// Topic counts too many for a per-wavefront LDS histogram (K above 40,960): each token adds its
// one to n_wk directly.
__global__ void k_count_direct(const CountTask* __restrict__ tasks, u32 n_tasks, const u16* __restrict__ zs,
                               u32* __restrict__ n, u32 K) {
  const u32 task = blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (task >= n_tasks) return;
  const CountTask tk = tasks[task];
  u32* row = n + (u64)tk.word * K;
  for (u32 p = lane_id(); p < tk.len; p += 8 * 64) {  // eight loads in flight, as in k_count
    u32 zb[8];
#pragma unroll
    for (u32 b = 0; b < 8; ++b) zb[b] = p + 64 * b < tk.len ? (u32)zs[tk.p0 + p + 64 * b] : 0u;
#pragma unroll
    for (u32 b = 0; b < 8; ++b)
      if (p + 64 * b < tk.len) atomicAdd(&row[zb[b]], 1u);
  }
}

// ---------------------------------------------------------------------------------------------
// Stage kernels used by the conformance tests (same device functions as the iteration).
// ---------------------------------------------------------------------------------------------
__global__ void k_stage_philox(const u32* ctr, const u32* key, u32* out, u32 n) {
  u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  u32x4 c;
  c.x = ctr[4 * i];
  c.y = ctr[4 * i + 1];
  c.z = ctr[4 * i + 2];
  c.w = ctr[4 * i + 3];
  u32x4 r = philox4x32_10(c, key[2 * i], key[2 * i + 1]);
  out[4 * i] = r.x;
  out[4 * i + 1] = r.y;
  out[4 * i + 2] = r.z;
  out[4 * i + 3] = r.w;
}

__global__ void k_stage_scan(const f32* in, f32* out, u32 n_groups) {
  u32 g = blockIdx.x;
  if (g >= n_groups) return;
  out[g * 64 + lane_id()] = wave_scan_f32(in[g * 64 + lane_id()]);
}

// ops[i] = key | (1u << 31) for a decrement (of a present key); the final map is written to out in
// entry form (key << 16 | count, kMapEmpty for an empty slot).  The operations are those of the
// sweep: a count change, and an insertion or a removal when the count leaves or reaches 0.
__global__ void k_stage_map(const u32* ops, u32 n_ops, u32 cap, u32 K, u32* out) {
  extern __shared__ u32 map_smem[];
  const u32 lane = lane_id();
  const DocLds m = doc_lds(map_smem, K);
  const u32 hm = map_home_mult(K);
  const u32 e2 = 2u * K;
  for (u32 i = lane; i < (K + 2) / 2; i += 64) ((u32*)m.C)[i] = 0u;
  for (u32 s = lane; s < cap; s += 64) m.Kt[s] = (u16)e2;
  for (u32 i = 0; i < n_ops; ++i) {
    const u32 op = ops[i];
    const u32 key2 = 2u * (op & 0x7FFFFFFFu);
    const bool dec = (op >> 31) != 0u;
    wave_sync();
    u32 old = count_at(m.C, key2);
    if (dec && old == 0u) continue;  // not a present key
    wave_sync();
    if (lane == 0) count_put(m.C, key2, dec ? old - 1u : old + 1u);
    wave_sync();
    if (!dec && old == 0u) keymap_insert(m.Kt, cap, hm, e2, key2);
    if (dec && old == 1u) keymap_erase(m.Kt, cap, hm, e2, key2);
  }
  wave_sync();
  for (u32 s = lane; s < cap; s += 64) {
    const u32 v = m.Kt[s];
    out[s] = v == e2 ? kMapEmpty : map_entry(v >> 1, count_at(m.C, v));
  }
}

// One wavefront per draw: the map `table` (cap slots in entry form, distinct keys) is loaded into
// the sweep's LDS state and draw i is made with the uniforms ux[i], uw[i] by the sweep's own device
// functions -- the register variants when use_regs (maps of up to eight lines), else the line-loop
// form.  out[i] is the picked slot, or kNoSlot for a dense draw.
__global__ __launch_bounds__(64) void k_stage_pick(const u32* table, u32 cap, u32 K, const f32* phi_row_k, f32 sa,
                                                   const f32* ux, const f32* uw, u32 count, u32 use_regs, u32* out) {
  extern __shared__ u32 pick_smem[];
  const u32 i = blockIdx.x;
  if (i >= count) return;
  const u32 lane = lane_id();
  const DocLds m = doc_lds(pick_smem, K);
  const u32 e2 = 2u * K;
  for (u32 q = lane; q < (K + 2) / 2; q += 64) ((u32*)m.C)[q] = 0u;
  wave_sync();
  for (u32 s = lane; s < cap; s += 64) {
    const u32 e = table[s];
    if (map_count(e) != 0u) {
      m.Kt[s] = (u16)(2u * map_key(e));
      count_put(m.C, 2u * map_key(e), map_count(e));
    } else {
      m.Kt[s] = (u16)e2;
    }
  }
  wave_sync();
  const u32 lines = uniform_u32(cap / kMapLine);
  const PhiRow prow = phi_row(phi_row_k, K);
  const TokenU u{sa, ux[i], uw[i], 0u};  // one token: the same values in every lane
  Draw d;
  if (use_regs && lines <= 8) {
    switch (lines) {
      case 1: d = draw_topic_regs_lds<1>(m, prow, u); break;
      case 2: d = draw_topic_regs_lds<2>(m, prow, u); break;
      case 3: d = draw_topic_regs_lds<3>(m, prow, u); break;
      case 4: d = draw_topic_regs_lds<4>(m, prow, u); break;
      case 5: d = draw_topic_regs_lds<5>(m, prow, u); break;
      case 6: d = draw_topic_regs_lds<6>(m, prow, u); break;
      case 7: d = draw_topic_regs_lds<7>(m, prow, u); break;
      default: d = draw_topic_regs_lds<8>(m, prow, u); break;
    }
  } else {
    d = draw_topic_loop(m.Kt, m.C, lines, prow, u, e2);
  }
  if (lane == 0) out[i] = d.key2 == kNoSlot ? kNoSlot : d.slot;
}

__global__ void k_stage_poisson(const u32* nval, u32 count, f32 beta, const AliasSlot* tables, u32 c1, u32 c2,
                                u32 tag, u32 k0, u32 k1, u32* out) {
  u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  out[i] = poisson_draw(nval[i], beta, tables, i, c1, c2, tag, k0, k1);
}

__global__ void k_stage_alias_draw(const AliasSlot* row, u32 m, const u32* r_slot, const u32* r_keep, u32 count,
                                   u32* out) {
  u32 i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  out[i] = draw_alias_slot(row, m, r_slot[i], r_keep[i]);
}

}  // namespace lda
