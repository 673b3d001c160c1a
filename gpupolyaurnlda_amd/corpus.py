"""Seeded synthetic LDA corpora (the five workload shapes the sampler is measured and checked on).

``generate(spec)`` returns word ids ``w`` (u32, document-major), document lengths ``d_len`` and
initial topics ``z`` (uniform over K, from the same Philox counters as the library's device
initialiser).  The generator is native (libgpulda_host.so, OpenMP) so that the 250M-token shape
takes seconds.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, replace

import numpy as np

from ._lib import GenSpec, hostlib, ptr

POISSON, LOGNORMAL, POWERLAW = 0, 1, 2


@dataclass(frozen=True)
class CorpusSpec:
    name: str
    V: int
    K: int
    D: int
    len_kind: int
    len_a: float
    len_b: float = 0.0
    len_c: float = 0.0
    gen_alpha: float = 0.1   # theta_d ~ Dirichlet(gen_alpha)
    gen_beta: float = 0.01   # phi_k ~ Dirichlet(gen_beta * zipf)
    zipf_s: float = 0.0
    seed: int = 0
    alpha: float = 0.1       # sampler prior
    beta: float = 0.01
    iterations: int = 1

    def scaled(self, D: int, name: str | None = None) -> "CorpusSpec":
        return replace(self, D=D, name=name or f"{self.name}[D={D}]")


# Plain sampler, tiny corpus (seed 0): V=1000, K=20, D=1000, L ~ Poisson(50); 100 iterations.
TINY = CorpusSpec("tiny V=1000 K=20 D=1000 L~Poisson(50)", 1000, 20, 1000, POISSON, 50.0,
                  gen_alpha=0.1, gen_beta=0.01, zipf_s=0.0, seed=0, alpha=0.1, beta=0.01, iterations=100)
# Medium corpus (seed 1): V=50k Zipf(1.07), K=256, D=200k, L ~ lognormal(100, 0.8) <= 2000; 200 iterations.
MEDIUM = CorpusSpec("medium V=50000 K=256 D=200000 L~LN(100,0.8)<=2000", 50000, 256, 200000, LOGNORMAL, 100.0,
                    0.8, 2000.0, gen_alpha=0.1, gen_beta=0.01, zipf_s=1.07, seed=1, alpha=0.1, beta=0.01,
                    iterations=200)
# Large corpus (seed 2): V=100k Zipf(1.07), K=1024, D=1M, L ~ lognormal(200, 1.0) <= 4096; alpha = 50/K.
LARGE = CorpusSpec("large V=100000 K=1024 D=1000000 L~LN(200,1.0)<=4096", 100000, 1024, 1000000, LOGNORMAL, 200.0,
                   1.0, 4096.0, gen_alpha=50.0 / 1024, gen_beta=0.01, zipf_s=1.07, seed=2, alpha=50.0 / 1024,
                   beta=0.01, iterations=10)
# Batched form (seed 3): V=100k, K=1024, D=2M, L ~ L^-1.8 on [4, 20000].
BATCHED = CorpusSpec("batched V=100000 K=1024 D=2000000 L~L^-1.8 on [4,20000]", 100000, 1024, 2000000, POWERLAW, 1.8,
                     4.0, 20000.0, gen_alpha=50.0 / 1024, gen_beta=0.01, zipf_s=1.07, seed=3, alpha=50.0 / 1024,
                     beta=0.01, iterations=10)

SPECS = {"tiny": TINY, "medium": MEDIUM, "large": LARGE, "batched": BATCHED}


def _genspec(s: CorpusSpec) -> GenSpec:
    return GenSpec(s.V, s.K, s.D, s.len_kind, s.len_a, s.len_b, s.len_c, s.gen_alpha, s.gen_beta, s.zipf_s, s.seed)


def lengths(spec: CorpusSpec) -> np.ndarray:
    d_len = np.empty(spec.D, dtype=np.uint32)
    gs = _genspec(spec)
    hostlib().gpulda_gen_lengths(C.byref(gs), ptr(d_len, C.c_uint32))
    return d_len


def generate(spec: CorpusSpec, with_z: bool = True, with_phi: bool = False):
    """Returns dict(w, d_len, z, doc_off, C[, phi_true])."""
    gs = _genspec(spec)
    d_len = np.empty(spec.D, dtype=np.uint32)
    N = int(hostlib().gpulda_gen_lengths(C.byref(gs), ptr(d_len, C.c_uint32)))
    w = np.empty(N, dtype=np.uint32)
    z = np.empty(N, dtype=np.uint32) if with_z else None
    phi = np.empty((spec.V, spec.K), dtype=np.float32) if with_phi else None
    hostlib().gpulda_gen_tokens(C.byref(gs), ptr(d_len, C.c_uint32), ptr(w, C.c_uint32), ptr(z, C.c_uint32),
                                ptr(phi, C.c_float))
    doc_off = np.zeros(spec.D + 1, dtype=np.uint64)
    np.cumsum(d_len, out=doc_off[1:])
    out = dict(w=w, d_len=d_len, z=z, doc_off=doc_off, C=np.bincount(w, minlength=spec.V).astype(np.uint32))
    if with_phi:
        out["phi_true"] = phi
    return out


def requests(phi: np.ndarray, n_req: int, median: float = 300.0, sigma: float = 1.2, lo: int = 10,
             hi: int = 5000, rate: float = 2000.0, seed: int = 4):
    """Fold-in request stream over a trained phi (V, K): (d_len, arrival seconds, w)."""
    phi = np.ascontiguousarray(phi, dtype=np.float32)
    V, K = phi.shape
    d_len = np.empty(n_req, dtype=np.uint32)
    arr = np.empty(n_req, dtype=np.float64)
    N = int(hostlib().gpulda_gen_requests(ptr(phi, C.c_float), V, K, n_req, median, sigma, lo, hi, rate, seed,
                                          ptr(d_len, C.c_uint32), ptr(arr, C.c_double), None))
    w = np.empty(N, dtype=np.uint32)
    hostlib().gpulda_gen_requests(ptr(phi, C.c_float), V, K, n_req, median, sigma, lo, hi, rate, seed,
                                  ptr(d_len, C.c_uint32), ptr(arr, C.c_double), ptr(w, C.c_uint32))
    return d_len, arr, w
