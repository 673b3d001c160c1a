"""Stage-level entry points of the HIP library on numpy arrays.

Each call stages device copies, runs the same device functions the iteration uses, and returns
host arrays.  They mirror the reference's per-kernel unit tests (src/cu/test/*.cu), which launch
one kernel on hand-made inputs.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib, ptr


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def philox(ctr, key) -> np.ndarray:
    """Philox4x32-10 blocks: ctr (n, 4) u32, key (n, 2) u32 -> (n, 4) u32."""
    ctr = _u32(ctr).reshape(-1, 4)
    key = _u32(key).reshape(-1, 2)
    out = np.empty_like(ctr)
    check(lib().gpulda_stage_philox(ptr(ctr, C.c_uint32), ptr(key, C.c_uint32), ptr(out, C.c_uint32), len(ctr)),
          "stage_philox")
    return out


def scan64(x) -> np.ndarray:
    """Canonical 64-lane inclusive f32 scan of each group of 64 values."""
    x = _f32(x).reshape(-1, 64)
    out = np.empty_like(x)
    check(lib().gpulda_stage_scan64(ptr(x, C.c_float), ptr(out, C.c_float), len(x)), "stage_scan64")
    return out


def map_apply(ops, cap: int, K: int) -> np.ndarray:
    """Applies increments (topic) / decrements (topic | 1<<31) to an empty doc-topic map of topics
    below K."""
    ops = _u32(ops)
    out = np.empty(cap, dtype=np.uint32)
    check(lib().gpulda_stage_map(ptr(ops, C.c_uint32), len(ops), cap, K, ptr(out, C.c_uint32)), "stage_map")
    return out


def alias(rows, divisor=None, alpha: float = 1.0):
    """Row normalisation, canonical row sums and alias tables.

    Returns (rows_out, sums, prob, alias)."""
    rows = _f32(rows)
    if rows.ndim == 1:
        rows = rows.reshape(1, -1)
    n_rows, m = rows.shape
    rows_out = np.empty_like(rows)
    sums = np.empty(n_rows, dtype=np.float32)
    prob = np.empty((n_rows, m), dtype=np.float32)
    al = np.empty((n_rows, m), dtype=np.uint32)
    div = None if divisor is None else np.ascontiguousarray(divisor, dtype=np.uint64)
    check(lib().gpulda_stage_alias(ptr(rows, C.c_float), n_rows, m, ptr(div, C.c_uint64), C.c_float(alpha),
                                   ptr(rows_out, C.c_float), ptr(sums, C.c_float), ptr(prob, C.c_float),
                                   ptr(al, C.c_uint32)), "stage_alias")
    return rows_out, sums, prob, al


def alias_draw(prob, al, r_slot, r_keep) -> np.ndarray:
    prob = _f32(prob)
    al = _u32(al)
    r_slot = _u32(r_slot)
    r_keep = _u32(r_keep)
    out = np.empty(len(r_slot), dtype=np.uint32)
    check(lib().gpulda_stage_alias_draw(ptr(prob, C.c_float), ptr(al, C.c_uint32), len(prob),
                                        ptr(r_slot, C.c_uint32), ptr(r_keep, C.c_uint32), len(r_slot),
                                        ptr(out, C.c_uint32)), "stage_alias_draw")
    return out


def pick(table, K: int, phi_row, sigma_a: float, u_decide, u_target, use_regs: bool = True) -> np.ndarray:
    """The sweep's sparse/dense decision and sparse pick on a given map: the picked slot for each
    (u_decide, u_target) pair, or -1 for a dense draw."""
    table = _u32(table)
    phi_row = _f32(phi_row)
    ux = _f32(np.atleast_1d(u_decide))
    uw = _f32(np.atleast_1d(u_target))
    assert len(ux) == len(uw) and len(phi_row) == K
    out = np.empty(len(ux), dtype=np.uint32)
    check(lib().gpulda_stage_pick(ptr(table, C.c_uint32), len(table), K, ptr(phi_row, C.c_float), C.c_float(sigma_a),
                                  ptr(ux, C.c_float), ptr(uw, C.c_float), len(ux), 1 if use_regs else 0,
                                  ptr(out, C.c_uint32)), "stage_pick")
    res = out.astype(np.int64)
    res[out == 0xFFFFFFFF] = -1
    return res


def doc_offsets(d_len) -> np.ndarray:
    """First token of each document (and the total) as the sweeps read them."""
    d_len = _u32(d_len)
    out = np.empty(len(d_len) + 1, dtype=np.uint64)
    check(lib().gpulda_stage_doc_offsets(ptr(d_len, C.c_uint32), len(d_len), ptr(out, C.c_uint64)),
          "stage_doc_offsets")
    return out


def poisson(n, beta: float, c1: int = 0, c2: int = 0, tag: int = 3, seed: int = 0) -> np.ndarray:
    """Poisson(n_i + beta) draws with Philox counters (i, c1, c2, tag)."""
    n = _u32(n)
    out = np.empty_like(n)
    check(lib().gpulda_stage_poisson(ptr(n, C.c_uint32), len(n), C.c_float(beta), c1, c2, tag, seed,
                                     ptr(out, C.c_uint32)), "stage_poisson")
    return out


def phi_draw(n, beta: float, iteration: int, seed: int = 0):
    """phi~ (V, K) f32 and exact topic sums (K,) u64 of the Polya-urn draw from counts n (V, K)."""
    n = _u32(n)
    V, K = n.shape
    phi = np.empty((V, K), dtype=np.float32)
    ts = np.empty(K, dtype=np.uint64)
    check(lib().gpulda_stage_phi_draw(ptr(n, C.c_uint32), V, K, C.c_float(beta), iteration, seed,
                                      ptr(phi, C.c_float), ptr(ts, C.c_uint64)), "stage_phi_draw")
    return phi, ts


def sweep(z, w, doc_off, doc_base: int, iteration: int, seed: int, K: int, phi, sigma_a, prob, al):
    """One z-sweep of every document; returns (new z, map entries visited)."""
    z = _u32(z).copy()
    w = _u32(w)
    doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
    phi = _f32(phi)
    V = phi.shape[0]
    sigma_a = _f32(sigma_a)
    prob = _f32(prob)
    al = _u32(al)
    visited = np.zeros(1, dtype=np.uint64)
    check(lib().gpulda_stage_sweep(ptr(z, C.c_uint32), ptr(w, C.c_uint32), ptr(doc_off, C.c_uint64),
                                   len(doc_off) - 1, doc_base, iteration, seed, K, V, ptr(phi, C.c_float),
                                   ptr(sigma_a, C.c_float), ptr(prob, C.c_float), ptr(al, C.c_uint32),
                                   ptr(visited, C.c_uint64)), "stage_sweep")
    return z, int(visited[0])


def count(w, z, V: int, K: int) -> np.ndarray:
    """n_wk (V, K) u32 by the word-sorted reduction."""
    w = _u32(w)
    z = _u32(z)
    out = np.empty((V, K), dtype=np.uint32)
    check(lib().gpulda_stage_count(ptr(w, C.c_uint32), ptr(z, C.c_uint32), len(w), V, K, ptr(out, C.c_uint32)),
          "stage_count")
    return out


def foldin(w, doc_off, req_ids, sweeps: int, seed: int, K: int, phi, sigma_a, prob, al):
    """Fold-in of each request (documents of w delimited by doc_off) against a fixed model;
    returns (topics, map entries visited)."""
    w = _u32(w)
    doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
    req_ids = _u32(req_ids)
    phi = _f32(phi)
    V = phi.shape[0]
    z = np.empty(len(w), dtype=np.uint32)
    visited = np.zeros(1, dtype=np.uint64)
    check(lib().gpulda_stage_foldin(ptr(w, C.c_uint32), ptr(doc_off, C.c_uint64), len(req_ids),
                                    ptr(req_ids, C.c_uint32), sweeps, seed, K, V, ptr(phi, C.c_float),
                                    ptr(_f32(sigma_a), C.c_float), ptr(_f32(prob), C.c_float),
                                    ptr(_u32(al), C.c_uint32), ptr(z, C.c_uint32), ptr(visited, C.c_uint64)),
          "stage_foldin")
    return z, int(visited[0])
