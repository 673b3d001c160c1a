"""Device-resident corpus: the whole corpus stays in HBM and iterations run without PCIe traffic.

This is the form used for throughput measurement: one ``iterate`` is exactly the reference's
iteration (sample_phi, then the z-sweep and n_wk reduction of every document; train.rs:85-102,
train.cu:137-172) with the corpus already on the device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Args, Buffer, check, lib, ptr


class ResidentCorpus:
    def __init__(self, w, d_len, K: int, V: int, alpha: float, beta: float, seed: int = 0, z=None,
                 counts_from_z: bool = True):
        self._lib = lib()
        self.K, self.V = int(K), int(V)
        self.w = np.ascontiguousarray(w, dtype=np.uint32)
        self.d_len = np.ascontiguousarray(d_len, dtype=np.uint32)
        self.n_tokens = int(self.w.size)
        self.n_docs = int(self.d_len.size)
        self.C = np.bincount(self.w, minlength=self.V).astype(np.uint32)
        self.args = Args(alpha, beta, self.K, self.V, ptr(self.C, C.c_uint32), 0, 0)
        self._buffers = (Buffer * 1)()
        check(self._lib.gpulda_set_seed(seed), "set_seed")
        self._lib.initialize(C.byref(self.args), self._buffers, 0)
        self._open = True
        zz = None if z is None else np.ascontiguousarray(z, dtype=np.uint32)
        check(self._lib.gpulda_corpus_load(ptr(self.w, C.c_uint32), ptr(self.d_len, C.c_uint32), self.n_docs,
                                           ptr(zz, C.c_uint32)), "corpus_load")
        if counts_from_z:
            check(self._lib.gpulda_corpus_counts_from_z(), "counts_from_z")

    def iterate(self, n: int = 1):
        """Runs n iterations; returns mean device ms per iteration of
        [Poisson draw, alias rows, z-sweep, n_wk reduction]."""
        ms = np.zeros(4, dtype=np.float32)
        check(self._lib.gpulda_corpus_iterate(n, ptr(ms, C.c_float)), "corpus_iterate")
        return ms

    BUCKETS, PERSISTENT, PERSISTENT_BLOCK = 0, 1, 2

    def schedule(self, kind: int, waves_per_cu: int = 16, block_min_len: int = 0) -> None:
        """Sweep schedule: BUCKETS (per-bucket launches), PERSISTENT (one queue-driven launch),
        PERSISTENT_BLOCK (plus block-per-document for documents >= block_min_len tokens)."""
        check(self._lib.gpulda_corpus_schedule(kind, waves_per_cu, block_min_len), "corpus_schedule")

    def use_graph(self, enabled: bool = True) -> None:
        check(self._lib.gpulda_corpus_use_graph(1 if enabled else 0), "corpus_use_graph")

    def autotune(self):
        """Times every candidate schedule on one sweep (chain unchanged) and applies the fastest.
        Returns (table rows {schedule, waves_per_cu, block_min_len, count_rows, ms}, chosen index)."""
        table = np.zeros((32, 5), dtype=np.float32)
        n = np.zeros(1, dtype=np.uint32)
        chosen = C.c_int(-1)
        check(self._lib.gpulda_corpus_autotune(ptr(table, C.c_float), 32, ptr(n, C.c_uint32), C.byref(chosen)),
              "corpus_autotune")
        rows = [dict(schedule=int(r[0]), waves_per_cu=int(r[1]), block_min_len=int(r[2]), count_rows=bool(r[3]),
                     ms=float(r[4])) for r in table[:int(n[0])]]
        return rows, int(chosen.value)

    def count_rows(self, enabled: bool = True) -> None:
        """u16 rows of the phi~ counts read by the persistent sweep where they fit (default on);
        z does not depend on it."""
        check(self._lib.gpulda_set_count_rows(1 if enabled else 0), "set_count_rows")

    def f32_rows(self, n_words: int) -> None:
        """The n_words most frequent words keep f32 rows (default: 1 MiB of rows)."""
        check(self._lib.gpulda_set_f32_rows(int(n_words)), "set_f32_rows")

    def check_count_rows(self):
        """(count rows on, cells whose phi recovered from the count rows differs from the f32 rows,
        words read as count rows) for the last Phi step."""
        on = C.c_int(0)
        bad = np.zeros(1, dtype=np.uint64)
        fit = np.zeros(1, dtype=np.uint64)
        check(self._lib.gpulda_check_count_rows(C.byref(on), ptr(bad, C.c_uint64), ptr(fit, C.c_uint64)),
              "check_count_rows")
        return bool(on.value), int(bad[0]), int(fit[0])

    def iterate_untimed(self, n: int = 1) -> None:
        check(self._lib.gpulda_corpus_iterate(n, None), "corpus_iterate")

    @property
    def iteration(self) -> int:
        return int(self._lib.gpulda_iteration())

    def z(self) -> np.ndarray:
        out = np.empty(self.n_tokens, dtype=np.uint32)
        check(self._lib.gpulda_corpus_copy_z(ptr(out, C.c_uint32)), "corpus_copy_z")
        return out

    def counts(self) -> np.ndarray:
        out = np.empty((self.V, self.K), dtype=np.uint32)
        check(self._lib.gpulda_copy_counts(ptr(out, C.c_uint32)), "copy_counts")
        return out

    def phi(self) -> np.ndarray:
        out = np.empty((self.V, self.K), dtype=np.float32)
        check(self._lib.gpulda_copy_phi(ptr(out, C.c_float)), "copy_phi")
        return out

    def sigma_a(self) -> np.ndarray:
        out = np.empty(self.V, dtype=np.float32)
        check(self._lib.gpulda_copy_sigma_a(ptr(out, C.c_float)), "copy_sigma_a")
        return out

    def alias(self):
        prob = np.empty((self.V, self.K), dtype=np.float32)
        al = np.empty((self.V, self.K), dtype=np.uint32)
        check(self._lib.gpulda_copy_alias(ptr(prob, C.c_float), ptr(al, C.c_uint32)), "copy_alias")
        return prob, al

    def topic_sums(self) -> np.ndarray:
        out = np.empty(self.K, dtype=np.uint64)
        check(self._lib.gpulda_copy_topic_sums(ptr(out, C.c_uint64)), "copy_topic_sums")
        return out

    def loglik(self) -> float:
        out = np.zeros(1, dtype=np.float64)
        check(self._lib.gpulda_corpus_loglik(ptr(out, C.c_double)), "corpus_loglik")
        return float(out[0])

    def visited(self) -> int:
        out = np.zeros(1, dtype=np.uint64)
        check(self._lib.gpulda_sweep_stats(ptr(out, C.c_uint64)), "sweep_stats")
        return int(out[0])

    def reset_stats(self) -> None:
        check(self._lib.gpulda_reset_stats(), "reset_stats")

    def close(self) -> None:
        if self._open:
            self._lib.gpulda_corpus_release()
            self._lib.cleanup(self._buffers, 0)
            self._open = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
