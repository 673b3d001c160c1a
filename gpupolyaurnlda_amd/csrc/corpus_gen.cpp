// corpus_gen.cpp -- seeded synthetic LDA corpora (host, OpenMP).
//
// A corpus is drawn from the LDA generative model:
//   phi_k   ~ Dirichlet(beta * z_v),  z_v = V (v+1)^-s / sum_u (u+1)^-s   (s = 0: flat)
//   theta_d ~ Dirichlet(alpha)
//   L_d     ~ Poisson(mean) | round(lognormal(median, sigma)) capped | discrete power law
//   token:  topic ~ theta_d, word ~ phi_topic
// Word ids are ranked by their Zipf weight (id 0 most frequent), as the reference's
// preprocess.rs:23 numbers its vocabulary by descending count.  Lengths are at least 1 because a
// zero length ends a buffer in train.rs:24,30.  Every random draw comes from Philox4x32-10 with a
// counter per (document or topic, draw index), so the output does not depend on the thread count.
// The initial topic indicators use the same counters as the library's device initialiser.
//
// Gamma variates: Marsaglia & Tsang, "A simple method for generating gamma variables" (ACM TOMS
// 2000), with the shape < 1 boost G(a) = G(a+1) U^(1/a) taken in log space.  Normals: Box-Muller.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "lda_core.h"

using namespace lda;

namespace {

enum GenTag : u32 { kGenLen = 0x41, kGenTheta = 0x42, kGenTok = 0x43, kGenPhi = 0x44, kGenReq = 0x45 };

struct Stream {
  u32 c0, c1, tag, k0, k1;
  u32 j = 0;
  u32 buf[4];
  int have = 0;
  Stream(u32 a, u32 b, u32 t, u64 seed) : c0(a), c1(b), tag(t), k0((u32)seed), k1((u32)(seed >> 32)) {}
  u32 next() {
    if (have == 0) {
      u32x4 c;
      c.x = c0;
      c.y = c1;
      c.z = j++;
      c.w = tag << 24;
      u32x4 r = philox4x32_10(c, k0, k1);
      buf[0] = r.x;
      buf[1] = r.y;
      buf[2] = r.z;
      buf[3] = r.w;
      have = 4;
    }
    return buf[--have];
  }
  double uniform() {  // (0, 1)
    u32 a = next(), b = next();
    return (((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) + 0.5) * 0x1p-53;
  }
  double normal() {
    double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  // log of a Gamma(a) variate
  double log_gamma_variate(double a) {
    if (a < 1.0) return log_gamma_variate(a + 1.0) + std::log(uniform()) / a;
    double d = a - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
    for (;;) {
      double x = normal();
      double v = 1.0 + c * x;
      if (v <= 0.0) continue;
      v = v * v * v;
      double u = uniform();
      if (std::log(u) < 0.5 * x * x + d - d * v + d * std::log(v)) return std::log(d * v);
    }
  }
};

// Dirichlet(shape_i) as a cumulative distribution (last entry 1).
void dirichlet_cdf(Stream& st, const double* shape, u32 n, double* cdf) {
  double mx = -INFINITY;
  for (u32 i = 0; i < n; ++i) {
    cdf[i] = st.log_gamma_variate(shape[i]);
    mx = std::max(mx, cdf[i]);
  }
  double s = 0.0;
  for (u32 i = 0; i < n; ++i) {
    s += std::exp(cdf[i] - mx);
    cdf[i] = s;
  }
  for (u32 i = 0; i < n; ++i) cdf[i] /= s;
  cdf[n - 1] = 1.0;
}

u32 cdf_pick(const double* cdf, u32 n, double u) {
  u32 i = (u32)(std::upper_bound(cdf, cdf + n, u) - cdf);
  return i < n ? i : n - 1;
}

}  // namespace

extern "C" {

struct GenSpec {
  uint32_t V, K, D;
  uint32_t len_kind;  // 0 Poisson(a), 1 lognormal(median a, sigma b, cap c), 2 power law (exponent a, lo b, hi c)
  double len_a, len_b, len_c;
  double alpha;
  double beta;
  double zipf_s;
  uint64_t seed;
};

// Phase 1: document lengths; returns the token count.
uint64_t gpulda_gen_lengths(const GenSpec* s, uint32_t* d_len) {
  const u32 D = s->D;
#pragma omp parallel for schedule(static)
  for (long d = 0; d < (long)D; ++d) {
    Stream st((u32)d, 0, kGenLen, s->seed);
    double L = 1;
    if (s->len_kind == 0) {  // Poisson by inversion
      double u = st.uniform(), p = std::exp(-s->len_a), c = p;
      u32 k = 0;
      while (u > c && k < 100000) {
        ++k;
        p *= s->len_a / k;
        c += p;
      }
      L = k;
    } else if (s->len_kind == 1) {
      L = std::floor(s->len_a * std::exp(s->len_b * st.normal()) + 0.5);
      L = std::min(L, s->len_c);
    } else {
      double a1 = 1.0 - s->len_a, lo = s->len_b, hi = s->len_c + 1.0;
      double u = st.uniform();
      L = std::floor(std::pow(std::pow(lo, a1) + u * (std::pow(hi, a1) - std::pow(lo, a1)), 1.0 / a1));
      L = std::min(std::max(L, s->len_b), s->len_c);
    }
    if (L < 1) L = 1;
    d_len[d] = (uint32_t)L;
  }
  uint64_t n = 0;
  for (u32 d = 0; d < D; ++d) n += d_len[d];
  return n;
}

// Topic-word distributions as column-stochastic phi (V x K word-major, f32) -- optional output.
// Phase 2: tokens.  w_out (and z_out, if not NULL) have sum(d_len) entries.
int gpulda_gen_tokens(const GenSpec* s, const uint32_t* d_len, uint32_t* w_out, uint32_t* z_out, float* phi_out) {
  const u32 V = s->V, K = s->K, D = s->D;
  std::vector<double> zipf(V);
  double zs = 0.0;
  for (u32 v = 0; v < V; ++v) {
    zipf[v] = s->zipf_s > 0 ? std::pow((double)v + 1.0, -s->zipf_s) : 1.0;
    zs += zipf[v];
  }
  for (u32 v = 0; v < V; ++v) zipf[v] = s->beta * zipf[v] * ((double)V / zs);
  std::vector<double> phi_cdf((u64)K * V);
#pragma omp parallel for schedule(dynamic, 1)
  for (long k = 0; k < (long)K; ++k) {
    Stream st((u32)k, 0, kGenPhi, s->seed);
    dirichlet_cdf(st, zipf.data(), V, &phi_cdf[(u64)k * V]);
  }
  if (phi_out) {
#pragma omp parallel for schedule(static)
    for (long v = 0; v < (long)V; ++v)
      for (u32 k = 0; k < K; ++k) {
        const double* c = &phi_cdf[(u64)k * V];
        phi_out[(u64)v * K + k] = (float)(c[v] - (v ? c[v - 1] : 0.0));
      }
  }
  std::vector<u64> off((u64)D + 1, 0);
  for (u32 d = 0; d < D; ++d) off[d + 1] = off[d] + d_len[d];
#pragma omp parallel
  {
    std::vector<double> theta(K), shape(K, s->alpha);
#pragma omp for schedule(dynamic, 64)
    for (long d = 0; d < (long)D; ++d) {
      Stream th((u32)d, 0, kGenTheta, s->seed);
      dirichlet_cdf(th, shape.data(), K, theta.data());
      Stream tk((u32)d, 1, kGenTok, s->seed);
      const u64 b = off[d];
      for (u32 i = 0; i < d_len[d]; ++i) {
        u32 topic = cdf_pick(theta.data(), K, tk.uniform());
        w_out[b + i] = cdf_pick(&phi_cdf[(u64)topic * V], V, tk.uniform());
        if (z_out) {
          u32x4 r = philox_tagged(i, (u32)d, 0, kTagInitZ, 0, (u32)s->seed, (u32)(s->seed >> 32));
          z_out[b + i] = bounded(r.x, K);
        }
      }
    }
  }
  return 0;
}

// Fold-in requests: n_req documents whose words come from a random mixture of 1..5 topics of a
// trained phi (V x K word-major, normalised over words per topic).  Lengths are
// round(lognormal(median, sigma)) clipped to [lo, hi]; arrival gaps are exponential at `rate`.
uint64_t gpulda_gen_requests(const float* phi, uint32_t V, uint32_t K, uint32_t n_req, double median, double sigma,
                             uint32_t lo, uint32_t hi, double rate, uint64_t seed, uint32_t* len_out,
                             double* arrival_out, uint32_t* w_out /* NULL: lengths and arrivals only */) {
  std::vector<u64> off((u64)n_req + 1, 0);
  double t = 0.0;
  for (u32 r = 0; r < n_req; ++r) {
    Stream st(r, 0, kGenReq, seed);
    double L = std::floor(median * std::exp(sigma * st.normal()) + 0.5);
    L = std::min<double>(std::max<double>(L, lo), hi);
    len_out[r] = (u32)L;
    t += -std::log(st.uniform()) / rate;
    arrival_out[r] = t;
    off[r + 1] = off[r] + len_out[r];
  }
  if (!w_out) return off[n_req];
  std::vector<double> cdf((u64)K * V);
#pragma omp parallel for schedule(dynamic, 1)
  for (long k = 0; k < (long)K; ++k) {
    double c = 0.0;
    for (u32 v = 0; v < V; ++v) {
      c += phi[(u64)v * K + k];
      cdf[(u64)k * V + v] = c;
    }
    for (u32 v = 0; v < V; ++v) cdf[(u64)k * V + v] = c > 0 ? cdf[(u64)k * V + v] / c : (double)(v + 1) / V;
    cdf[(u64)k * V + V - 1] = 1.0;
  }
#pragma omp parallel for schedule(dynamic, 16)
  for (long r = 0; r < (long)n_req; ++r) {
    Stream st((u32)r, 1, kGenReq, seed);
    u32 m = 1 + (u32)(st.uniform() * 5.0);
    if (m > 5) m = 5;
    u32 topics[5];
    double wts[5], ws = 0.0;
    for (u32 j = 0; j < m; ++j) {
      topics[j] = (u32)(st.uniform() * K) % K;
      wts[j] = -std::log(st.uniform());
      ws += wts[j];
    }
    for (u32 i = 0; i < len_out[r]; ++i) {
      double u = st.uniform() * ws, c = 0.0;
      u32 j = 0;
      for (; j + 1 < m; ++j) {
        c += wts[j];
        if (u < c) break;
      }
      w_out[off[r] + i] = cdf_pick(&cdf[(u64)topics[j] * V], V, st.uniform());
    }
  }
  return off[n_req];
}

}  // extern "C"
