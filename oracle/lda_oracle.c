/*
 * lda_oracle.c -- CPU restatement of the Polya-urn LDA iteration, used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only to check or time against the HIP path.  Nothing in the
 * product (gpupolyaurnlda_amd/) links, loads or calls it.
 *
 * It is written independently of the product sources (plain sequential C, textbook data
 * structures) and follows the same arithmetic specification, so that agreement between the two
 * is evidence for both.  Each function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).
 *
 * Specification summary (one Gibbs iteration t, seed s, key = (s_lo, s_hi)):
 *   Phi step     x_vk = PoissonDraw(n_vk + beta; Philox(k, v, t, PHI))      polyaurn.cu:87-110
 *                S_k  = sum_v x_vk (exact, 64-bit integer)                  polyaurn.cu:113
 *                Phi_vk = (f32)x_vk / (f32)S_k                              polyaurn.cu:116-122
 *   colsums      T_v = canonical 64-lane sum of Phi_v.;  sigma_a[v] = alpha*T_v  polyaurn.cu:144-168
 *   alias        exact fixed-point split of Phi_v. / T_v                   polyaurn.cu:171-177,
 *                                                                          spalias.cu:88-174
 *   z sweep      per document, sequential over tokens:                    topics.cu:36-105
 *                remove z; r = Philox(i, doc, t, SWEEP); if T_w == 0 keep z; else
 *                sigma_b = grouped column sums of count*Phi_wk over map slots, then one
 *                canonical scan; sparse if u(r.x)*(sigma_a+sigma_b) > sigma_a: the slot whose
 *                prefix interval holds u(r.w)*sigma_b (column, then slot in the column),
 *                else alias draw (r.y slot, r.z keep); add the new z.
 *   counts       n_vk = #tokens with (w = v, z = k)                         topics.cu:95, polyaurn.cu:133-141
 *
 * Parity status of each function is recorded in the file header of tests/test_oracle.py.
 */
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint8_t u8;
typedef uint16_t u16;
typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

/* ------------------------------------------------------------------------------------------ */
/* Philox4x32-10 (SC'11).  Replaces cuRAND Philox4_32_10 of random.cu:6-17.                   */
/* ------------------------------------------------------------------------------------------ */
void orc_philox(const u32 ctr[4], const u32 key[2], u32 out[4]) {
  u32 x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
  u32 k0 = key[0], k1 = key[1];
  for (int round = 0; round < 10; ++round) {
    u64 a = (u64)0xD2511F53u * x0;
    u64 b = (u64)0xCD9E8D57u * x2;
    u32 y0 = (u32)(b >> 32) ^ x1 ^ k0;
    u32 y1 = (u32)b;
    u32 y2 = (u32)(a >> 32) ^ x3 ^ k1;
    u32 y3 = (u32)a;
    x0 = y0; x1 = y1; x2 = y2; x3 = y3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

enum { TAG_INIT_Z = 1, TAG_INIT_N = 2, TAG_PHI = 3, TAG_SWEEP = 4, TAG_FOLD_INIT = 5, TAG_FOLD_SWEEP = 6 };

static void draw4(u32 c0, u32 c1, u32 c2, u32 tag, u32 sub, u64 seed, u32 out[4]) {
  u32 ctr[4] = {c0, c1, c2, (tag << 24) | (sub & 0xFFFFFFu)};
  u32 key[2] = {(u32)seed, (u32)(seed >> 32)};
  orc_philox(ctr, key, out);
}

static float unif24(u32 r) { return (float)(r >> 8) / 16777216.0f; }
static double unif53(u32 a, u32 b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) / 9007199254740992.0;
}
static u32 below(u32 r, u32 n) { return (u32)(((u64)r * n) >> 32); }

/* ------------------------------------------------------------------------------------------ */
/* Deterministic double helpers (log by atanh series, sqrt by Newton, log k! by Stirling).     */
/* Needed because the PTRS accept test must take the same decision on every machine.          */
/* ------------------------------------------------------------------------------------------ */
double orc_log(double x) {
  u64 bits;
  memcpy(&bits, &x, 8);
  int ex = (int)((bits >> 52) & 0x7FF);
  if (ex == 0) {
    x *= 18014398509481984.0; /* 2^54 */
    memcpy(&bits, &x, 8);
    ex = (int)((bits >> 52) & 0x7FF) - 54;
  }
  ex -= 1023;
  u64 mb = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
  double m;
  memcpy(&m, &mb, 8);
  if (m > 1.4142135623730951) { m *= 0.5; ex += 1; }
  double f = (m - 1.0) / (m + 1.0);
  double f2 = f * f;
  /* atanh(f)/f = sum_j f^(2j)/(2j+1), j = 0..10, Horner from the top */
  double p = 1.0 / 21.0;
  for (int d = 19; d >= 1; d -= 2) p = p * f2 + 1.0 / (double)d;
  double lm = (2.0 * f) * p;
  double e = (double)ex;
  return e * 6.93147180369123816490e-01 + (lm + e * 1.90821492927058770002e-10);
}

double orc_sqrt(double x) {
  u64 bits;
  memcpy(&bits, &x, 8);
  bits = (bits >> 1) + 0x1FF8000000000000ull;
  double y;
  memcpy(&y, &bits, 8);
  for (int i = 0; i < 6; ++i) y = 0.5 * (y + x / y);
  return y;
}

double orc_log_factorial(i64 k) {
  if (k < 16) {
    double s = 0.0;
    for (i64 i = 2; i <= k; ++i) s = s + orc_log((double)i);
    return s;
  }
  double x = (double)(k + 1);
  double ix = 1.0 / x;
  double ix2 = ix * ix;
  double corr = ix * (1.0 / 12.0 - ix2 * (1.0 / 360.0 - ix2 * (1.0 / 1260.0)));
  return (x - 0.5) * orc_log(x) - x + 0.91893853320467274178 + corr;
}

/* Hoermann (1993) PTRS.  Replaces the Gaussian branch of draw_poisson, polyaurn.cu:30-35. */
u32 orc_poisson_ptrs(double lam, u32 c0, u32 c1, u32 c2, u32 tag, u64 seed) {
  double slam = orc_sqrt(lam);
  double loglam = orc_log(lam);
  double b = 0.931 + 2.53 * slam;
  double a = -0.059 + 0.02483 * b;
  double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
  double vr = 0.9277 - 3.6224 / (b - 2.0);
  for (u32 att = 0; att < 256; ++att) {
    u32 r[4];
    draw4(c0, c1, c2, tag, att, seed, r);
    double U = unif53(r[0], r[1]) - 0.5;
    double V = unif53(r[2], r[3]);
    double us = 0.5 - fabs(U);
    if (us <= 0.0) continue;
    double kf = (2.0 * a / us + b) * U + lam + 0.43;
    if (kf < 0.0 || kf >= 4294967295.0) continue;
    i64 k = (i64)kf;
    if (us >= 0.07 && V <= vr) return (u32)k;
    if (us < 0.013 && V > us) continue;
    if (V <= 0.0) return (u32)k;
    double lhs = orc_log(V) + orc_log(inv_alpha) - orc_log(a / (us * us) + b);
    double rhs = -lam + (double)k * loglam - orc_log_factorial(k);
    if (lhs <= rhs) return (u32)k;
  }
  return (u32)(lam + 0.5);
}

/* ------------------------------------------------------------------------------------------ */
/* Canonical 64-lane f32 networks (shared definition with the device kernels).                 */
/* scan: Hillis-Steele inside each 16-lane row (offsets 1,2,4,8), then rows 1 and 3 add the    */
/*       last value of rows 0 and 2, then rows 2 and 3 add lane 31 (T1 + T0): row 3 ends as    */
/*       (x + T2) + (T1 + T0) (T_r = row r total).                                             */
/* ------------------------------------------------------------------------------------------ */
void orc_scan64(float v[64]) {
  for (int off = 1; off < 16; off <<= 1) {
    float nv[64];
    for (int l = 0; l < 64; ++l) nv[l] = (l % 16 >= off) ? v[l] + v[l - off] : v[l];
    memcpy(v, nv, sizeof(nv));
  }
  /* rows 1 and 3 add the last value of rows 0 and 2; then rows 2 and 3 add lane 31 */
  float t0 = v[15], t2 = v[47];
  for (int l = 16; l < 32; ++l) v[l] = v[l] + t0;
  for (int l = 48; l < 64; ++l) v[l] = v[l] + t2;
  float c = v[31];
  for (int l = 32; l < 64; ++l) v[l] = v[l] + c;
}

/* Sum of x[0..n): lane l accumulates x[l], x[l+64], ... left to right, then scan64's total. */
float orc_sum64(const float* x, u32 n) {
  float acc[64];
  for (int l = 0; l < 64; ++l) {
    float a = 0.0f;
    for (u64 i = (u64)l; i < n; i += 64) a = a + x[i];
    acc[l] = a;
  }
  orc_scan64(acc);
  return acc[63];
}

/* ------------------------------------------------------------------------------------------ */
/* Alias table, exact split in 2^-32 fixed point.  Replaces build_alias, spalias.cu:88-174     */
/* (same Vose pairing of under-full slots with over-full ones; the pairing is computed from    */
/* integer prefix sums, so it is unique instead of queue-order dependent).                     */
/* ------------------------------------------------------------------------------------------ */
typedef struct { float prob; u32 alias; } OrcAlias;

void orc_alias_build(const float* w, float total, u32 n, OrcAlias* out) {
  if (!(total > 0.0f)) {
    for (u32 i = 0; i < n; ++i) { out[i].prob = 1.0f; out[i].alias = i; }
    return;
  }
  /* fixed-point weights q_i = floor(w_i / total * n * 2^32); slot capacity Q = sum q, item
   * weight a_i = n * q_i, so deficits and surpluses balance exactly */
  u64* a = (u64*)malloc(sizeof(u64) * n);
  u32* heavy = (u32*)malloc(sizeof(u32) * n);
  u64* E = (u64*)malloc(sizeof(u64) * n);
  u64* D = (u64*)malloc(sizeof(u64) * n);
  u64 Q = 0;
  for (u32 i = 0; i < n; ++i) {
    double ratio = (double)w[i] / (double)total;
    a[i] = (u64)(ratio * ((double)n * 4294967296.0));
    Q += a[i];
  }
  for (u32 i = 0; i < n; ++i) a[i] *= n;
  u32 nh = 0, nl = 0;
  u64 se = 0, sd = 0;
  for (u32 i = 0; i < n; ++i) {
    if (a[i] >= Q) { se += a[i] - Q; heavy[nh] = i; E[nh] = se; ++nh; }
    else { D[nl] = sd; sd += Q - a[i]; ++nl; }
  }
  u32 li = 0, hi = 0;
  for (u32 i = 0; i < n; ++i) {
    if (a[i] < Q) {
      u64 d = D[li++];
      u32 j = 0; /* first heavy with E > d */
      while (j < nh && E[j] <= d) ++j;
      if (j == nh) j = nh - 1;
      out[i].prob = (float)((double)a[i] / (double)Q);
      out[i].alias = heavy[j];
    } else {
      u32 j = hi++;
      u64 e = E[j];
      u32 b = 0; /* first light with D >= e; its D (or the total deficit) ends j's turn */
      while (b < nl && D[b] < e) ++b;
      u64 dend = b < nl ? D[b] : sd;
      if (dend > e && j + 1 < nh) {
        out[i].prob = (float)((double)(Q + e - dend) / (double)Q);
        out[i].alias = heavy[j + 1];
      } else {
        out[i].prob = 1.0f;
        out[i].alias = i;
      }
    }
  }
  free(a); free(heavy); free(E); free(D);
}

/* draw_alias (topics.cuh:63-83) with the slot and the keep test on two independent words: slot =
 * floor(r_slot * n / 2^32), kept when u(r_keep) < prob[slot], else its alias. */
static u32 draw_alias(const OrcAlias* row, u32 n, u32 r_slot, u32 r_keep) {
  u32 slot = below(r_slot, n);
  return unif24(r_keep) < row[slot].prob ? slot : row[slot].alias;
}

u32 orc_alias_draw(const OrcAlias* row, u32 n, u32 r_slot, u32 r_keep) { return draw_alias(row, n, r_slot, r_keep); }

/* Poisson tables.  Restates Poisson::Poisson + build_poisson_prob, poisson.cu:9-34: table n is */
/* the PMF of Poisson(n + beta) over 0..199 (computed in double, stored as f32).               */
void orc_poisson_tables(float beta, OrcAlias* tables /* 100*200 */) {
  float pmf[200];
  for (u32 n = 0; n < 100; ++n) {
    double lam = (double)n + (double)beta;
    for (u32 x = 0; x < 200; ++x)
      pmf[x] = (float)exp((double)x * log(lam) - lam - lgamma((double)x + 1.0));
    orc_alias_build(pmf, orc_sum64(pmf, 200), 200, tables + (u64)n * 200);
  }
}

/* draw_poisson, polyaurn.cu:7-36 (table for n < 100, PTRS above). */
u32 orc_poisson(u32 n, float beta, const OrcAlias* tables, u32 c0, u32 c1, u32 c2, u32 tag, u64 seed) {
  if (n < 100) {
    u32 r[4];
    draw4(c0, c1, c2, tag, 0, seed, r);
    return draw_alias(tables + (u64)n * 200, 200, r[0], r[1]);
  }
  return orc_poisson_ptrs((double)n + (double)beta, c0, c1, c2, tag, seed);
}

/* ------------------------------------------------------------------------------------------ */
/* Phi step.  Restates polya_urn_sample (polyaurn.cu:87-123) with word-major arrays [v*K+k].   */
/* ------------------------------------------------------------------------------------------ */
void orc_phi_draw(const u32* n, u32 V, u32 K, float beta, u32 iter, u64 seed, const OrcAlias* tables,
                  float* phi_tilde, u64* topic_sum) {
  for (u32 k = 0; k < K; ++k) topic_sum[k] = 0;
#ifdef _OPENMP
#pragma omp parallel
#endif
  {
    u64* part = (u64*)calloc(K, sizeof(u64));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (long v = 0; v < (long)V; ++v)
      for (u32 k = 0; k < K; ++k) {
        u32 x = orc_poisson(n[(u64)v * K + k], beta, tables, k, (u32)v, iter, TAG_PHI, seed);
        phi_tilde[(u64)v * K + k] = (float)x;
        part[k] += x;
      }
#ifdef _OPENMP
#pragma omp critical
#endif
    for (u32 k = 0; k < K; ++k) topic_sum[k] += part[k]; /* integer sums: order-free */
    free(part);
  }
}

/* polya_urn_init (polyaurn.cu:61-79): n_vk ~ Pois(floor(C_v / K) + beta), as integers. */
void orc_init_counts(const u32* C, u32 V, u32 K, float beta, u64 seed, const OrcAlias* tables, u32* n) {
  for (u32 v = 0; v < V; ++v)
    for (u32 k = 0; k < K; ++k)
      n[(u64)v * K + k] = orc_poisson(C[v] / K, beta, tables, k, v, 0, TAG_INIT_N, seed);
}

/* Normalise, colsums and alias: polya_urn_sample's divide (polyaurn.cu:116-122),
 * polya_urn_colsums (polyaurn.cu:144-178) and build_alias (spalias.cu:88-174), per word. */
void orc_phi_finish(float* phi /* in: phi~, out: phi */, const u64* topic_sum, u32 V, u32 K, float alpha,
                    float* sigma_a, OrcAlias* alias) {
  float* sf = (float*)malloc(sizeof(float) * K);
  for (u32 k = 0; k < K; ++k) sf[k] = (float)(double)topic_sum[k];
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (long v = 0; v < (long)V; ++v) {
    float* row = phi + (u64)v * K;
    for (u32 k = 0; k < K; ++k) row[k] = sf[k] > 0.0f ? row[k] / sf[k] : 0.0f;
    float t = orc_sum64(row, K);
    sigma_a[v] = alpha * t;
    orc_alias_build(row, t, K, alias + (u64)v * K);
  }
  free(sf);
}

/* ------------------------------------------------------------------------------------------ */
/* Doc-topic map: textbook sequential Robin Hood linear probing, ties broken by key, backward  */
/* shift deletion.  Restates the HashMap of hashmap.cuh:13-334 (entry = key<<16 | count).     */
/* ------------------------------------------------------------------------------------------ */
/* A stored entry has count >= 1; every empty slot holds count 0 at key 0xFFFF (its mass 0 * phi
 * is +0 whatever the key). */
#define IS_EMPTY(e) (((e) & 0xFFFFu) == 0u)
#define EMPTY_SLOT 0xFFFF0000u

u32 orc_map_capacity(u32 L, u32 K) {
  u32 need = L < K ? L : K;
  u32 c = need + need / 32 + 1;
  return ((c + 63) / 64) * 64;
}
/* home slot of a topic below K: key * hm, hm = floor((2^32 - 1) / K), scaled to [0, cap) by
 * multiply-shift -- non-decreasing in the key, so the table is ordered by topic */
u32 orc_map_home_mult(u32 K) { return 0xFFFFFFFFu / K; }
static u32 home_of(u32 key, u32 cap, u32 hm) { return (u32)(((u64)(key * hm) * cap) >> 32); }
static u32 dist(u32 s, u32 h, u32 cap) { return (s + cap - h) % cap; }

/* returns slot of key or -1 */
long orc_map_find(const u32* T, u32 cap, u32 hm, u32 key) {
  u32 h = home_of(key, cap, hm);
  for (u32 d = 0; d < cap; ++d) {
    u32 s = (h + d) % cap;
    u32 e = T[s];
    if (IS_EMPTY(e)) return -1;
    if ((e >> 16) == key) return (long)s;
    u32 de = dist(s, home_of(e >> 16, cap, hm), cap);
    if (de < d || (de == d && (e >> 16) > key)) return -1;
  }
  return -1;
}

void orc_map_inc(u32* T, u32 cap, u32 hm, u32 key) {
  long s = orc_map_find(T, cap, hm, key);
  if (s >= 0) { T[s] += 1; return; }
  u32 cur = (key << 16) | 1u;
  u32 slot = home_of(key, cap, hm);
  u32 d = 0;
  for (;;) {
    u32 e = T[slot];
    if (IS_EMPTY(e)) { T[slot] = cur; return; }
    u32 de = dist(slot, home_of(e >> 16, cap, hm), cap);
    if (de < d || (de == d && (e >> 16) > (cur >> 16))) {
      T[slot] = cur;
      cur = e;
      d = de;
    }
    slot = (slot + 1) % cap;
    d += 1;
  }
}

/* Inserts an absent key with its count (Robin Hood order). */
static void map_put(u32* T, u32 cap, u32 hm, u32 cur) {
  u32 slot = home_of(cur >> 16, cap, hm);
  u32 d = 0;
  for (;;) {
    u32 e = T[slot];
    if (IS_EMPTY(e)) { T[slot] = cur; return; }
    u32 de = dist(slot, home_of(e >> 16, cap, hm), cap);
    if (de < d || (de == d && (e >> 16) > (cur >> 16))) {
      T[slot] = cur;
      cur = e;
      d = de;
    }
    slot = (slot + 1) % cap;
    d += 1;
  }
}

/* Moves every entry of a table of capacity `cap` into a cleared table of capacity `ncap`
 * (same array, ncap >= cap). */
static void map_regrow(u32* T, u32 cap, u32 ncap, u32 K, u32* scratch) {
  const u32 hm = orc_map_home_mult(K);
  u32 n = 0;
  for (u32 s = 0; s < cap; ++s)
    if (!IS_EMPTY(T[s])) scratch[n++] = T[s];
  for (u32 s = 0; s < ncap; ++s) T[s] = EMPTY_SLOT;
  for (u32 i = 0; i < n; ++i) map_put(T, ncap, hm, scratch[i]);
}

void orc_map_dec(u32* T, u32 cap, u32 hm, u32 key) {
  long s = orc_map_find(T, cap, hm, key);
  if (s < 0) abort();
  if ((T[s] & 0xFFFFu) > 1) { T[s] -= 1; return; }
  u32 hole = (u32)s;
  for (;;) {
    u32 nx = (hole + 1) % cap;
    u32 e = T[nx];
    if (IS_EMPTY(e) || dist(nx, home_of(e >> 16, cap, hm), cap) == 0) { T[hole] = EMPTY_SLOT; return; }
    T[hole] = e;
    hole = nx;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* z sweep of one document.  Restates sample_topics (topics.cu:36-105) with count_topics      */
/* (topics.cuh:142-158), compute_product_cumsum (topics.cuh:163-192, with Phi[w][k] -- the     */
/* reference reads Phi[k] only), draw_wary_search (topics.cuh:88-137: the slot whose prefix    */
/* interval holds the target), draw_alias (topics.cuh:63-83).                                  */
/* Summation order of the sparse mass (the reference's is its warp scan; this one is fixed so   */
/* the device can evaluate it with one wave scan per token):                                   */
/*   m[j][l] = count * phi[w][key] of slot j*64+l (0 when empty);                              */
/*   G_g[l]  = m[g][l] + m[g+4][l] + ... left to right (line group g = j % 4);                 */
/*   A[l]    = ((G_0 + G_1) + G_2) + G_3;  P = scan64(A);  sigma_b = P[63].                    */
/* The sparse pick for target t: lane l* = first l with A[l] > 0 and P[l] > t (else the last   */
/* with A[l] > 0); E = P[l*-1] (0 for l* = 0); Q = scan64 of column l*'s masses m[j][l*] in  */
/* line order (64 lines per scan, carry = running total of earlier scans); the line is the     */
/* first with m > 0 and E + (carry + Q_j) > t (else the last line with m > 0).                 */
/* Returns the number of map entries that the sparse sums visited (sum over tokens).           */
/* ------------------------------------------------------------------------------------------ */
/* Capacity rule: the map starts sized for the document's distinct topics n0
 * (orc_map_capacity(n0, K)) and, when a new topic would exceed its load
 * (orc_map_capacity(keys + 1, K) > cap), grows to orc_map_capacity(min(2 (keys + 1), min(L, K)), K).
 * The reference starts from the K_d hint and regrows by 1.2x (hashmap.cuh:98, 159-207). */
typedef struct {
  u32* T;
  float* mass;
  u32* scratch;
  u32 L, K, cap, keys, kmax, hm;
} OrcDoc;

/* Opens a document: scratch holds >= 5 * orc_map_capacity(L, K) words. */
static OrcDoc doc_open(const u32* z, u32 L, u32 K, u32* scratch) {
  OrcDoc d;
  const u32 cap_max = orc_map_capacity(L, K);
  d.T = scratch;
  d.mass = (float*)(d.T + cap_max);
  d.scratch = (u32*)(d.mass + cap_max);
  d.L = L;
  d.K = K;
  d.hm = orc_map_home_mult(K);
  d.kmax = L < K ? L : K;
  u32 n0 = 0;
  for (u32 s = 0; s < cap_max; ++s) d.T[s] = EMPTY_SLOT;
  for (u32 i = 0; i < L; ++i) {
    if (orc_map_find(d.T, cap_max, d.hm, z[i]) < 0) ++n0;
    orc_map_inc(d.T, cap_max, d.hm, z[i]);
  }
  d.cap = orc_map_capacity(n0, K);
  map_regrow(d.T, cap_max, d.cap, K, d.scratch); /* the same key set at the starting capacity */
  d.keys = n0;
  return d;
}

/* Sparse/dense decision and sparse pick for one token over a map of `lines` 64-slot lines (slot
 * j*64+l is line j, lane l) and the Phi row prow of the token's word (sigma_a = sa > 0):
 *   m[j][l] = count * prow[key] (0 when empty), written to mass[];
 *   G_g[l]  = m[g][l] + m[g+4][l] + ... left to right (line group g = j % 4);
 *   A[l]    = ((G_0 + G_1) + G_2) + G_3;  P = scan64(A);  sigma_b = P[63];
 *   sparse iff ux * (sa + sigma_b) > sa; then t = uw * sigma_b and the pick is the slot whose
 *   interval holds t when the slots are laid end to end column by column (lane 0's lines in line
 *   order, then lane 1's, ...): draw_wary_search (topics.cuh:88-137) restated two-level.
 *   lane l* = first l with A[l] > 0 and P[l] > t (else the last with A[l] > 0); E = P[l*-1]
 *   (0 for l* = 0); Q = scan64 of column l*'s masses in line order (64 lines per scan, carry =
 *   running total of earlier scans); the line is the first with m > 0 and E + (carry + Q_j) > t
 *   (else the last line with m > 0).
 * Returns the picked slot, or -1 for a dense draw (or a map without mass).  *visited, when not
 * NULL, is increased by the number of occupied slots the sum read. */
long orc_sparse_pick(const u32* T, u32 lines, const float* prow, float sa, float ux, float uw, float* mass,
                     u64* visited) {
  float G[4][64];
  for (int g = 0; g < 4; ++g)
    for (int l = 0; l < 64; ++l) G[g][l] = 0.0f;
  for (u32 j = 0; j < lines; ++j) {
    for (u32 l = 0; l < 64; ++l) {
      u32 e = T[j * 64 + l];
      if (visited) *visited += !IS_EMPTY(e);
      float m = IS_EMPTY(e) ? 0.0f : (float)(e & 0xFFFFu) * prow[e >> 16];
      mass[j * 64 + l] = m;
      G[j % 4][l] = G[j % 4][l] + m;
    }
  }
  float A[64], P[64];
  for (int l = 0; l < 64; ++l) {
    A[l] = ((G[0][l] + G[1][l]) + G[2][l]) + G[3][l];
    P[l] = A[l];
  }
  orc_scan64(P);
  float sb = P[63];
  if (!(ux * (sa + sb) > sa)) return -1;
  float t = uw * sb;
  long ls = -1, lpos = -1;
  for (int l = 0; l < 64; ++l) {
    if (A[l] > 0.0f) {
      lpos = l;
      if (ls < 0 && P[l] > t) ls = l;
    }
  }
  if (ls < 0) ls = lpos;
  if (ls < 0) return -1;
  /* in column l*: canonical scans of its masses in line order, 64 lines at a time */
  float E = ls > 0 ? P[ls - 1] : 0.0f;
  float carry = 0.0f;
  long hit = -1, last = -1;
  for (u32 c0 = 0; c0 < lines && hit < 0; c0 += 64) {
    float Q[64];
    for (u32 k = 0; k < 64; ++k) Q[k] = c0 + k < lines ? mass[(c0 + k) * 64 + (u32)ls] : 0.0f;
    float mc[64];
    memcpy(mc, Q, sizeof(Q));
    orc_scan64(Q);
    for (u32 k = 0; k < 64 && c0 + k < lines; ++k) {
      if (mc[k] > 0.0f) {
        if (hit < 0 && E + (carry + Q[k]) > t) hit = (long)(c0 + k);
        last = (long)(c0 + k);
      }
    }
    carry = carry + Q[63];
  }
  long js = hit >= 0 ? hit : last;
  return js * 64 + ls;
}

/* One pass over the document's tokens with counters (token, doc, iter, tag); returns the number
 * of map entries the sparse sums visited. */
static u64 doc_pass(OrcDoc* d, u32* z, const u32* w, u32 doc, u32 iter, u32 tag, u64 seed, const float* phi,
                    const float* sigma_a, const OrcAlias* alias) {
  const u32 K = d->K;
  u64 visited = 0;
  for (u32 i = 0; i < d->L; ++i) {
    u32 wi = w[i], zi = z[i];
    u32 r[4];
    draw4(i, doc, iter, tag, 0, seed, r);
    long at = orc_map_find(d->T, d->cap, d->hm, zi);
    if ((d->T[at] & 0xFFFFu) == 1) --d->keys;
    orc_map_dec(d->T, d->cap, d->hm, zi);
    float sa = sigma_a[wi];
    u32 znew = zi;
    if (sa > 0.0f) {
      long s = orc_sparse_pick(d->T, d->cap / 64, phi + (u64)wi * K, sa, unif24(r[0]), unif24(r[3]), d->mass,
                               &visited);
      if (s >= 0) znew = d->T[s] >> 16;
      else znew = draw_alias(alias + (u64)wi * K, K, r[1], r[2]);
    }
    if (orc_map_find(d->T, d->cap, d->hm, znew) < 0) {
      if (orc_map_capacity(d->keys + 1, K) > d->cap) {
        u32 want = 2 * (d->keys + 1) < d->kmax ? 2 * (d->keys + 1) : d->kmax;
        u32 ncap = orc_map_capacity(want, K);
        map_regrow(d->T, d->cap, ncap, K, d->scratch);
        d->cap = ncap;
      }
      ++d->keys;
    }
    orc_map_inc(d->T, d->cap, d->hm, znew);
    z[i] = znew;
  }
  return visited;
}

u64 orc_sweep_doc(u32* z, const u32* w, u32 L, u32 doc, u32 iter, u64 seed, u32 K, const float* phi,
                  const float* sigma_a, const OrcAlias* alias, u32* T /* scratch >= 5*capacity(L,K) */) {
  OrcDoc d = doc_open(z, L, K, T);
  return doc_pass(&d, z, w, doc, iter, TAG_SWEEP, seed, phi, sigma_a, alias);
}

/* Fold-in of one request against a fixed model: topics start from the word's alias table
 * (a draw from p(k | w) proportional to phi[w][k]; uniform for an all-zero row), then `sweeps`
 * passes with counters (token, request, sweep s = 1..sweeps, FOLD_SWEEP), the map carried from
 * pass to pass and shrunk before a pass when its keys need less capacity. */
u64 orc_foldin_doc(u32* z, const u32* w, u32 L, u32 req, u32 sweeps, u64 seed, u32 K, const float* phi,
                   const float* sigma_a, const OrcAlias* alias, u32* scratch) {
  for (u32 i = 0; i < L; ++i) {
    u32 r[4];
    draw4(i, req, 0, TAG_FOLD_INIT, 0, seed, r);
    z[i] = sigma_a[w[i]] > 0.0f ? draw_alias(alias + (u64)w[i] * K, K, r[1], r[2]) : below(r[0], K);
  }
  OrcDoc d = doc_open(z, L, K, scratch);
  u64 visited = 0;
  for (u32 s = 1; s <= sweeps; ++s) {
    /* the map is carried from pass to pass; once the document's topics have thinned out it is
     * rebuilt at the capacity its keys need, orc_map_capacity(keys, K) */
    if (s > 1 && orc_map_capacity(d.keys, K) < d.cap) {
      u32 ncap = orc_map_capacity(d.keys, K);
      map_regrow(d.T, d.cap, ncap, K, d.scratch);
      d.cap = ncap;
    }
    visited += doc_pass(&d, z, w, req, s, TAG_FOLD_SWEEP, seed, phi, sigma_a, alias);
  }
  return visited;
}

u64 orc_foldin(u32* z, const u32* w, const u64* off, u32 n_req, const u32* req_ids, u32 sweeps, u64 seed, u32 K,
               const float* phi, const float* sigma_a, const OrcAlias* alias, int threads) {
  u32 maxL = 0;
  for (u32 r = 0; r < n_req; ++r) {
    u32 L = (u32)(off[r + 1] - off[r]);
    if (L > maxL) maxL = L;
  }
  u32 need = 5 * orc_map_capacity(maxL, K);
  u64 visited = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads > 0 ? threads : 1) reduction(+ : visited)
#endif
  {
    u32* T = (u32*)malloc(sizeof(u32) * need);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (long r = 0; r < (long)n_req; ++r)
      visited += orc_foldin_doc(z + off[r], w + off[r], (u32)(off[r + 1] - off[r]), req_ids[r], sweeps, seed, K, phi,
                                sigma_a, alias, T);
    free(T);
  }
  (void)threads;
  return visited;
}

/* sample_z over a whole buffer; documents are independent (one OpenMP task each). */
u64 orc_sweep(u32* z, const u32* w, const u64* doc_off, u32 n_docs, u32 doc_base, u32 iter, u64 seed,
              u32 K, const float* phi, const float* sigma_a, const OrcAlias* alias, int threads) {
  u32 maxL = 0;
  for (u32 d = 0; d < n_docs; ++d) {
    u32 L = (u32)(doc_off[d + 1] - doc_off[d]);
    if (L > maxL) maxL = L;
  }
  u32 maxcap = 5 * orc_map_capacity(maxL, K);
  u64 visited = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads > 0 ? threads : 1) reduction(+ : visited)
#endif
  {
    u32* T = (u32*)malloc(sizeof(u32) * (maxcap > 0 ? maxcap : 64));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
    for (long d = 0; d < (long)n_docs; ++d) {
      u64 b = doc_off[d];
      u32 L = (u32)(doc_off[d + 1] - b);
      visited += orc_sweep_doc(z + b, w + b, L, doc_base + (u32)d, iter, seed, K, phi, sigma_a, alias, T);
    }
    free(T);
  }
  (void)threads;
  return visited;
}

/* n_wk as a brute-force histogram (the reduction that topics.cu:95's atomics perform). */
void orc_count(const u32* w, const u32* z, u64 N, u32 K, u32* n, u64 nsize) {
  memset(n, 0, sizeof(u32) * nsize);
  for (u64 i = 0; i < N; ++i) n[(u64)w[i] * K + z[i]] += 1;
}

/* Threads OpenMP will use for the parallel loops above. */
int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Map layout after building from a list of keys (below K) with increments (+1) and decrements
 * (-1), for the table-layout tests. ops[i] = key | (is_dec << 31). */
void orc_map_apply(u32* T, u32 cap, u32 K, const u32* ops, u32 n_ops) {
  const u32 hm = orc_map_home_mult(K);
  for (u32 s = 0; s < cap; ++s) T[s] = EMPTY_SLOT;
  for (u32 i = 0; i < n_ops; ++i) {
    if (ops[i] >> 31) orc_map_dec(T, cap, hm, ops[i] & 0x7FFFFFFFu);
    else orc_map_inc(T, cap, hm, ops[i]);
  }
}

/* Collapsed joint log-likelihood log p(w, z | alpha, beta) in double (Griffiths & Steyvers). */
double orc_loglik(const u32* n_wk, u32 V, u32 K, const u32* z, const u64* doc_off, u32 n_docs,
                  double alpha, double beta) {
  double ll = 0.0;
  double* nk = (double*)calloc(K, sizeof(double));
  for (u32 v = 0; v < V; ++v)
    for (u32 k = 0; k < K; ++k) {
      u32 c = n_wk[(u64)v * K + k];
      nk[k] += c;
      if (c) ll += lgamma(beta + c) - lgamma(beta);
    }
  for (u32 k = 0; k < K; ++k) ll += lgamma(V * beta) - lgamma(V * beta + nk[k]);
  u32* cnt = (u32*)calloc(K, sizeof(u32));
  for (u32 d = 0; d < n_docs; ++d) {
    u64 b = doc_off[d], e = doc_off[d + 1];
    for (u64 i = b; i < e; ++i) cnt[z[i]]++;
    ll += lgamma(K * alpha) - lgamma(K * alpha + (double)(e - b));
    for (u64 i = b; i < e; ++i)
      if (cnt[z[i]]) { ll += lgamma(alpha + cnt[z[i]]) - lgamma(alpha); cnt[z[i]] = 0; }
  }
  free(nk);
  free(cnt);
  return ll;
}
