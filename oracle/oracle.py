"""ctypes wrapper of the CPU oracle (oracle/liblda_oracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  The product package never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liblda_oracle.so")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f32p = C.POINTER(C.c_float)


class OrcAlias(C.Structure):
    _fields_ = [("prob", C.c_float), ("alias", C.c_uint32)]


_SIGS = {
    "orc_philox": (None, [u32p, u32p, u32p]),
    "orc_log": (C.c_double, [C.c_double]),
    "orc_sqrt": (C.c_double, [C.c_double]),
    "orc_log_factorial": (C.c_double, [C.c_int64]),
    "orc_poisson_ptrs": (C.c_uint32, [C.c_double, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64]),
    "orc_scan64": (None, [f32p]),
    "orc_sum64": (C.c_float, [f32p, C.c_uint32]),
    "orc_alias_build": (None, [f32p, C.c_float, C.c_uint32, C.POINTER(OrcAlias)]),
    "orc_poisson_tables": (None, [C.c_float, C.POINTER(OrcAlias)]),
    "orc_poisson": (C.c_uint32, [C.c_uint32, C.c_float, C.POINTER(OrcAlias), C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.c_uint32, C.c_uint64]),
    "orc_phi_draw": (None, [u32p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.c_uint64, C.POINTER(OrcAlias),
                            f32p, u64p]),
    "orc_init_counts": (None, [u32p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, C.POINTER(OrcAlias), u32p]),
    "orc_phi_finish": (None, [f32p, u64p, C.c_uint32, C.c_uint32, C.c_float, f32p, C.POINTER(OrcAlias)]),
    "orc_map_capacity": (C.c_uint32, [C.c_uint32, C.c_uint32]),
    "orc_map_apply": (None, [u32p, C.c_uint32, C.c_uint32, u32p, C.c_uint32]),
    "orc_map_home_mult": (C.c_uint32, [C.c_uint32]),
    "orc_sweep": (C.c_uint64, [u32p, u32p, u64p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, f32p,
                               f32p, C.POINTER(OrcAlias), C.c_int]),
    "orc_foldin": (C.c_uint64, [u32p, u32p, u64p, C.c_uint32, u32p, C.c_uint32, C.c_uint64, C.c_uint32, f32p,
                                f32p, C.POINTER(OrcAlias), C.c_int]),
    "orc_count": (None, [u32p, u32p, C.c_uint64, C.c_uint32, u32p, C.c_uint64]),
    "orc_threads": (C.c_int, []),
    "orc_loglik": (C.c_double, [u32p, C.c_uint32, C.c_uint32, u32p, u64p, C.c_uint32, C.c_double, C.c_double]),
    "orc_sparse_pick": (C.c_long, [u32p, C.c_uint32, f32p, C.c_float, C.c_float, C.c_float, f32p, u64p]),
    "orc_alias_draw": (C.c_uint32, [C.POINTER(OrcAlias), C.c_uint32, C.c_uint32, C.c_uint32]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            from . import build as _b  # type: ignore  # noqa: F401
        _lib = C.CDLL(SO)
        for name, (res, argt) in _SIGS.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = argt
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def philox(ctr, key):
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, dtype=np.uint32).reshape(-1, 2)
    out = np.empty_like(ctr)
    L = lib()
    for i in range(len(ctr)):
        L.orc_philox(_p(ctr[i], C.c_uint32), _p(key[i], C.c_uint32), _p(out[i], C.c_uint32))
    return out


def scan64(x):
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, 64).copy()
    for g in range(len(x)):
        lib().orc_scan64(_p(x[g], C.c_float))
    return x


def sum64(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return np.float32(lib().orc_sum64(_p(x, C.c_float), len(x)))


def _alias_arrays(arr):
    a = np.ctypeslib.as_array(arr)
    return a["prob"].copy(), a["alias"].copy()


def alias_build(w, total=None):
    w = np.ascontiguousarray(w, dtype=np.float32)
    if total is None:
        total = sum64(w)
    out = (OrcAlias * len(w))()
    lib().orc_alias_build(_p(w, C.c_float), C.c_float(total), len(w), out)
    return _alias_arrays(out)


def poisson_tables(beta):
    out = (OrcAlias * 20000)()
    lib().orc_poisson_tables(C.c_float(beta), out)
    prob, al = _alias_arrays(out)
    return prob.reshape(100, 200), al.reshape(100, 200), out


def poisson(n, beta, c1=0, c2=0, tag=3, seed=0, tables=None):
    if tables is None:
        tables = poisson_tables(beta)[2]
    n = np.asarray(n, dtype=np.uint32)
    return np.array([lib().orc_poisson(int(v), C.c_float(beta), tables, i, c1, c2, tag, seed) for i, v in enumerate(n)],
                    dtype=np.uint32)


def sparse_pick(T, prow, sa, ux, uw):
    """Sparse/dense decision and sparse pick of one token over map T (len a multiple of 64) and the
    Phi row of its word: the picked slot for each (ux, uw) pair, or -1 for a dense draw."""
    T = np.ascontiguousarray(T, dtype=np.uint32)
    prow = np.ascontiguousarray(prow, dtype=np.float32)
    ux, uw = np.broadcast_arrays(np.atleast_1d(np.asarray(ux, dtype=np.float32)),
                                 np.atleast_1d(np.asarray(uw, dtype=np.float32)))
    assert len(T) % 64 == 0
    mass = np.empty(len(T), dtype=np.float32)
    L = lib()
    return np.array([L.orc_sparse_pick(_p(T, C.c_uint32), len(T) // 64, _p(prow, C.c_float), C.c_float(sa),
                                       C.c_float(float(a)), C.c_float(float(b)), _p(mass, C.c_float), None)
                     for a, b in zip(ux, uw)], dtype=np.int64)


def alias_draw(prob, al, r_slot, r_keep):
    packed = _AliasPacked(np.asarray(prob, dtype=np.float32), np.asarray(al, dtype=np.uint32))
    n = len(np.asarray(prob).reshape(-1))
    return np.array([lib().orc_alias_draw(packed.arr, n, int(a), int(b)) for a, b in zip(r_slot, r_keep)],
                    dtype=np.uint32)


def map_capacity(L, K):
    return int(lib().orc_map_capacity(L, K))


def map_home_mult(K):
    return int(lib().orc_map_home_mult(K))


def map_apply(ops, cap, K):
    ops = np.ascontiguousarray(ops, dtype=np.uint32)
    T = np.empty(cap, dtype=np.uint32)
    lib().orc_map_apply(_p(T, C.c_uint32), cap, K, _p(ops, C.c_uint32), len(ops))
    return T


def phi_draw(n, beta, iteration, seed, tables=None):
    n = np.ascontiguousarray(n, dtype=np.uint32)
    V, K = n.shape
    if tables is None:
        tables = poisson_tables(beta)[2]
    phi = np.empty((V, K), dtype=np.float32)
    ts = np.empty(K, dtype=np.uint64)
    lib().orc_phi_draw(_p(n, C.c_uint32), V, K, C.c_float(beta), iteration, seed, tables, _p(phi, C.c_float),
                       _p(ts, C.c_uint64))
    return phi, ts


def init_counts(Cw, K, beta, seed, tables=None):
    Cw = np.ascontiguousarray(Cw, dtype=np.uint32)
    V = len(Cw)
    if tables is None:
        tables = poisson_tables(beta)[2]
    n = np.empty((V, K), dtype=np.uint32)
    lib().orc_init_counts(_p(Cw, C.c_uint32), V, K, C.c_float(beta), seed, tables, _p(n, C.c_uint32))
    return n


def phi_finish(phi_tilde, topic_sum, alpha):
    phi = np.ascontiguousarray(phi_tilde, dtype=np.float32).copy()
    V, K = phi.shape
    ts = np.ascontiguousarray(topic_sum, dtype=np.uint64)
    sig = np.empty(V, dtype=np.float32)
    out = (OrcAlias * (V * K))()
    lib().orc_phi_finish(_p(phi, C.c_float), _p(ts, C.c_uint64), V, K, C.c_float(alpha), _p(sig, C.c_float), out)
    prob, al = _alias_arrays(out)
    return phi, sig, prob.reshape(V, K), al.reshape(V, K)


class _AliasPacked:
    """Packs (prob, alias) into the oracle's struct array once."""

    def __init__(self, prob, al):
        n = prob.size
        self.arr = (OrcAlias * n)()
        a = np.ctypeslib.as_array(self.arr)
        a["prob"] = prob.reshape(-1)
        a["alias"] = al.reshape(-1)


def sweep(z, w, doc_off, doc_base, iteration, seed, K, phi, sigma_a, prob, al, threads=1):
    z = np.ascontiguousarray(z, dtype=np.uint32).copy()
    w = np.ascontiguousarray(w, dtype=np.uint32)
    doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
    phi = np.ascontiguousarray(phi, dtype=np.float32)
    sigma_a = np.ascontiguousarray(sigma_a, dtype=np.float32)
    packed = _AliasPacked(np.asarray(prob, dtype=np.float32), np.asarray(al, dtype=np.uint32))
    visited = lib().orc_sweep(_p(z, C.c_uint32), _p(w, C.c_uint32), _p(doc_off, C.c_uint64), len(doc_off) - 1,
                              doc_base, iteration, seed, K, _p(phi, C.c_float), _p(sigma_a, C.c_float),
                              packed.arr, threads)
    return z, int(visited)


def foldin(w, doc_off, req_ids, sweeps, seed, K, phi, sigma_a, prob, al, threads=1):
    """Fold-in of each request (documents of w delimited by doc_off) against a fixed model."""
    w = np.ascontiguousarray(w, dtype=np.uint32)
    doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
    req_ids = np.ascontiguousarray(req_ids, dtype=np.uint32)
    z = np.zeros(len(w), dtype=np.uint32)
    phi = np.ascontiguousarray(phi, dtype=np.float32)
    sigma_a = np.ascontiguousarray(sigma_a, dtype=np.float32)
    packed = _AliasPacked(np.asarray(prob, dtype=np.float32), np.asarray(al, dtype=np.uint32))
    visited = lib().orc_foldin(_p(z, C.c_uint32), _p(w, C.c_uint32), _p(doc_off, C.c_uint64), len(req_ids),
                               _p(req_ids, C.c_uint32), sweeps, seed, K, _p(phi, C.c_float),
                               _p(sigma_a, C.c_float), packed.arr, threads)
    return z, int(visited)


def count(w, z, V, K):
    w = np.ascontiguousarray(w, dtype=np.uint32)
    z = np.ascontiguousarray(z, dtype=np.uint32)
    n = np.empty((V, K), dtype=np.uint32)
    lib().orc_count(_p(w, C.c_uint32), _p(z, C.c_uint32), len(w), K, _p(n, C.c_uint32), V * K)
    return n


def loglik(n_wk, z, doc_off, alpha, beta):
    n_wk = np.ascontiguousarray(n_wk, dtype=np.uint32)
    V, K = n_wk.shape
    z = np.ascontiguousarray(z, dtype=np.uint32)
    doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
    return float(lib().orc_loglik(_p(n_wk, C.c_uint32), V, K, _p(z, C.c_uint32), _p(doc_off, C.c_uint64),
                                  len(doc_off) - 1, alpha, beta))


class Model:
    """A full CPU iteration sequence mirroring the library's state machine.

    counts: 'C' (initialize(): n ~ Pois(floor(C_v/K) + beta)) or 'z' (counts of the initial z)."""

    def __init__(self, w, doc_off, z, V, K, alpha, beta, seed=0, counts="z", threads=1):
        self.w = np.ascontiguousarray(w, dtype=np.uint32)
        self.doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
        self.z = np.ascontiguousarray(z, dtype=np.uint32).copy()
        self.V, self.K, self.alpha, self.beta, self.seed = V, K, float(np.float32(alpha)), float(np.float32(beta)), seed
        self.threads = threads
        self.tables = poisson_tables(beta)[2]
        if counts == "z":
            self.n = count(self.w, self.z, V, K)
        else:
            Cw = np.bincount(self.w, minlength=V).astype(np.uint32)
            self.n = init_counts(Cw, K, beta, seed, self.tables)
        self.iteration = 0
        self.visited = 0

    def phi_step(self):
        self.iteration += 1
        pt, self.topic_sum = phi_draw(self.n, self.beta, self.iteration, self.seed, self.tables)
        self.phi, self.sigma_a, self.prob, self.al = phi_finish(pt, self.topic_sum, self.alpha)

    def z_step(self, doc_base=0, lo=0, hi=None):
        hi = len(self.doc_off) - 1 if hi is None else hi
        b, e = int(self.doc_off[lo]), int(self.doc_off[hi])
        off = self.doc_off[lo:hi + 1] - self.doc_off[lo]
        znew, vis = sweep(self.z[b:e], self.w[b:e], off, doc_base, self.iteration, self.seed, self.K, self.phi,
                          self.sigma_a, self.prob, self.al, self.threads)
        self.z[b:e] = znew
        self.visited += vis

    def iterate(self):
        self.phi_step()
        self.z_step()
        self.n = count(self.w, self.z, self.V, self.K)
