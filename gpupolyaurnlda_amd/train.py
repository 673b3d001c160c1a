"""Host-side mirror of the reference's training driver (src/train.rs) over the five-call ABI.

``Sampler`` owns the ``Args`` and ``Buffer`` structs for the lifetime of the library state, as
train.rs does (train.rs:63-73, 104).  ``train`` replays train.rs:85-102: per iteration one
``sample_phi``, then buffers packed by ``fill_buffer`` (train.rs:19-45) are swept with
``sample_z_async`` + ``sync_buffer`` and the cursors advanced by ``empty_buffer`` (train.rs:47-58).

One deliberate difference: train.rs treats a zero document length as end-of-file
(``read_u32().unwrap_or(0)``, train.rs:24,30), so an empty document silently ends the iteration.
Here end-of-file is the end of the length array and empty documents are passed through.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import Args, Buffer, check, lib, ptr


class Sampler:
    """The library's global state, driven through initialize / sample_phi / sample_z_async /
    sync_buffer / cleanup (src/ffi.rs:15-61)."""

    def __init__(self, alpha: float, beta: float, K: int, C_counts, buffer_size: int, max_D: int,
                 n_buffers: int = 1, seed: int = 0):
        self._lib = lib()
        self.C = np.ascontiguousarray(C_counts, dtype=np.uint32)
        self.K = int(K)
        self.V = int(len(self.C))
        self.args = Args(alpha, beta, self.K, self.V, ptr(self.C, C.c_uint32), buffer_size, max_D)
        self.buffers = (Buffer * n_buffers)()
        self._views = [None] * n_buffers  # keep the numpy slices alive while the library reads them
        check(self._lib.gpulda_set_seed(seed), "set_seed")
        self._lib.initialize(C.byref(self.args), self.buffers, n_buffers)
        self._open = True

    # -- the five calls ------------------------------------------------------------------------
    def sample_phi(self) -> None:
        self._lib.sample_phi()

    def set_buffer(self, i: int, z: np.ndarray, w: np.ndarray, d: np.ndarray, k_d: np.ndarray) -> None:
        """Points buffer i at host slices (Buffer::set_z/set_w/set_d/set_k_d, buffer.rs:32-62)."""
        b = self.buffers[i]
        assert z.dtype == np.uint32 and w.dtype == np.uint32 and d.dtype == np.uint32 and k_d.dtype == np.uint32
        self._views[i] = (z, w, d, k_d)
        b.z = ptr(z, C.c_uint32)
        b.w = ptr(w, C.c_uint32)
        b.d = ptr(d, C.c_uint32)
        b.K_d = ptr(k_d, C.c_uint32)
        b.n_docs = len(d)
        b.n_tokens = len(w)

    def sample_z_async(self, i: int = 0) -> None:
        self._lib.sample_z_async(C.byref(self.buffers[i]))

    def sync_buffer(self, i: int = 0) -> None:
        self._lib.sync_buffer(C.byref(self.buffers[i]))

    def cleanup(self) -> None:
        if self._open:
            self._lib.cleanup(self.buffers, len(self.buffers))
            self._open = False

    # -- read-back (extensions) ----------------------------------------------------------------
    @property
    def iteration(self) -> int:
        return int(self._lib.gpulda_iteration())

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

    def counts(self) -> np.ndarray:
        out = np.empty((self.V, self.K), dtype=np.uint32)
        check(self._lib.gpulda_copy_counts(ptr(out, C.c_uint32)), "copy_counts")
        return out

    def set_counts(self, n: np.ndarray) -> None:
        n = np.ascontiguousarray(n, dtype=np.uint32)
        assert n.shape == (self.V, self.K)
        check(self._lib.gpulda_set_counts(ptr(n, C.c_uint32)), "set_counts")

    def topic_sums(self) -> np.ndarray:
        out = np.empty(self.K, dtype=np.uint64)
        check(self._lib.gpulda_copy_topic_sums(ptr(out, C.c_uint64)), "copy_topic_sums")
        return out

    def poisson_tables(self):
        prob = np.empty((100, 200), dtype=np.float32)
        al = np.empty((100, 200), dtype=np.uint32)
        check(self._lib.gpulda_copy_poisson_tables(ptr(prob, C.c_float), ptr(al, C.c_uint32)), "copy_tables")
        return prob, al

    def visited(self) -> int:
        out = np.zeros(1, dtype=np.uint64)
        check(self._lib.gpulda_sweep_stats(ptr(out, C.c_uint64)), "sweep_stats")
        return int(out[0])

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.cleanup()


def fill_buffer(d: np.ndarray, doc_pos: int, buffer_size: int, max_D: int):
    """Documents and tokens of the next buffer starting at document doc_pos (train.rs:19-45).

    Documents are taken while the buffer has room for one more document (max_D) and its tokens
    (buffer_size).  Returns (n_docs, n_tokens).  Raises if a document it examines exceeds
    buffer_size (train.rs:27-29): every document taken and the one that ends the buffer."""
    window = np.asarray(d[doc_pos:doc_pos + max_D + 1], dtype=np.uint64)
    if window.size == 0:
        return 0, 0
    cs = np.cumsum(window[:max_D])
    n_docs = int(np.searchsorted(cs, buffer_size, side="right"))
    examined = window[:n_docs + 1]
    if int(examined.max()) > buffer_size:
        first = int(np.argmax(examined > buffer_size))
        raise ValueError(f"document size exceeded buffer size: {int(examined[first])}")
    n_tokens = int(cs[n_docs - 1]) if n_docs else 0
    return n_docs, n_tokens


def train(sampler: Sampler, z: np.ndarray, w: np.ndarray, d: np.ndarray, k_d: np.ndarray, n_mc: int,
          buffer_size: int, max_D: int, on_iteration=None) -> None:
    """Replays train.rs:85-102 (z is updated in place, as the reference's z.bin mmap)."""
    for it in range(n_mc):
        sampler.sample_phi()
        doc_pos = 0
        tok_pos = 0
        while True:
            n_docs, n_tokens = fill_buffer(d, doc_pos, buffer_size, max_D)
            if n_docs == 0:
                break
            sampler.set_buffer(0, z[tok_pos:tok_pos + n_tokens], w[tok_pos:tok_pos + n_tokens],
                               d[doc_pos:doc_pos + n_docs], k_d[doc_pos:doc_pos + n_docs])
            sampler.sample_z_async(0)
            sampler.sync_buffer(0)
            doc_pos += n_docs
            tok_pos += n_tokens
        if on_iteration is not None:
            on_iteration(it)
