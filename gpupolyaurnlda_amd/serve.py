"""Fold-in serving over the resident model (phi, sigma_a and alias tables stay in HBM).

``FoldInServer`` wraps the library's continuous-batching server (gpulda_serve_*).  ``run_stream``
plays an open-loop request stream against it: requests are admitted as their arrival time passes
(and as slots and HBM arena space free up), admitted requests are published to the device's
request ring every ``window`` seconds -- where the resident worker pool picks each up at once,
beside the requests already running -- and each result is collected as soon as its completion
flag is raised.  Latency is completion minus arrival.  ``run_serial`` is the baseline: the same
requests one at a time, each submitted alone and waited for before the next.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from ._lib import check, lib, ptr


class FoldInServer:
    def __init__(self, hbm_budget_bytes: int = 16 << 30, sweeps: int = 50, waves_per_cu: int = 2,
                 max_slots: int = 16384):
        self._lib = lib()
        check(self._lib.gpulda_serve_open(hbm_budget_bytes, sweeps, waves_per_cu, max_slots), "serve_open")
        self._done = np.zeros(max_slots, dtype=np.uint32)
        self._n = np.zeros(1, dtype=np.uint32)
        self.open = True

    def submit(self, w: np.ndarray, request_id: int):
        """Slot of the admitted request, or None when there is no room now."""
        w = np.ascontiguousarray(w, dtype=np.uint32)
        slot = C.c_uint32(0)
        rc = self._lib.gpulda_serve_submit(ptr(w, C.c_uint32), len(w), request_id, C.byref(slot))
        if rc == 1:
            return None
        check(rc, "serve_submit")
        return int(slot.value)

    def flush(self) -> None:
        check(self._lib.gpulda_serve_flush(), "serve_flush")

    def poll(self):
        check(self._lib.gpulda_serve_poll(ptr(self._done, C.c_uint32), len(self._done), ptr(self._n, C.c_uint32)),
              "serve_poll")
        return self._done[:int(self._n[0])].tolist()

    def result(self, slot: int, length: int):
        z = np.empty(length, dtype=np.uint32)
        rid = C.c_uint32(0)
        check(self._lib.gpulda_serve_result(slot, ptr(z, C.c_uint32), C.byref(rid)), "serve_result")
        return z, int(rid.value)

    def close(self) -> None:
        if self.open:
            self._lib.gpulda_serve_close()
            self.open = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def foldin_noalloc(w, doc_off, req_ids, sweeps: int):
    """Fold-in of each request (documents of w delimited by doc_off) against the library's current
    model, through the allocation-free entry point.  Returns (topics, map entries visited)."""
    w = np.ascontiguousarray(w, dtype=np.uint32)
    doc_off = np.ascontiguousarray(doc_off, dtype=np.uint64)
    req_ids = np.ascontiguousarray(req_ids, dtype=np.uint32)
    z = np.empty(len(w), dtype=np.uint32)
    visited = np.zeros(1, dtype=np.uint64)
    check(lib().gpulda_foldin_noalloc(ptr(w, C.c_uint32), ptr(doc_off, C.c_uint64), len(req_ids),
                                      ptr(req_ids, C.c_uint32), sweeps, ptr(z, C.c_uint32),
                                      ptr(visited, C.c_uint64)), "foldin_noalloc")
    return z, int(visited[0])


def _split(d_len, w):
    off = np.zeros(len(d_len) + 1, dtype=np.int64)
    np.cumsum(d_len, out=off[1:])
    return [w[off[i]:off[i + 1]] for i in range(len(d_len))]


def _stalled(last_progress: float, limit: float) -> None:
    if time.perf_counter() - last_progress > limit:
        raise RuntimeError(f"fold-in server made no progress for {limit:.0f} s")


def run_stream(server: FoldInServer, d_len, arrival, w, window: float = 0.0002, id_base: int = 0,
               max_batch: int = 4096, stall_s: float = 60.0):
    """Open-loop replay of a request stream.  Returns dict(latency_s, in_server_s (admission to
    completion), results, wall_s)."""
    docs = _split(d_len, w)
    n = len(docs)
    latency = np.zeros(n)
    in_server = np.zeros(n)  # admission to completion
    admitted_at = np.zeros(n)
    results = [None] * n
    slot_of = {}
    nxt = 0
    done = 0
    pending = 0
    last_flush = 0.0
    t0 = time.perf_counter()
    progress = t0
    while done < n:
        _stalled(progress, stall_s)
        now = time.perf_counter() - t0
        while nxt < n and arrival[nxt] <= now and pending < max_batch:
            slot = server.submit(docs[nxt], id_base + nxt)
            if slot is None:
                break
            slot_of[slot] = nxt
            admitted_at[nxt] = time.perf_counter() - t0
            nxt += 1
            pending += 1
        if pending and (now - last_flush >= window or pending >= max_batch):
            server.flush()
            pending = 0
            last_flush = now
        for slot in server.poll():
            progress = time.perf_counter()
            i = slot_of.pop(slot)
            t_done = time.perf_counter() - t0
            latency[i] = t_done - arrival[i]
            in_server[i] = t_done - admitted_at[i]
            results[i], _ = server.result(slot, len(docs[i]))
            done += 1
        if pending == 0 and nxt < n and done < nxt:
            # nothing to launch: wait for the next arrival or completion without burning a core
            wait = arrival[nxt] - (time.perf_counter() - t0)
            if wait > 2e-4:
                time.sleep(min(wait, 2e-4))
    return {"latency_s": latency, "in_server_s": in_server, "results": results, "wall_s": time.perf_counter() - t0}


def run_serial(server: FoldInServer, d_len, w, id_base: int = 0, stall_s: float = 60.0):
    """Each request submitted alone and awaited before the next.  Returns dict(service_s, results)."""
    docs = _split(d_len, w)
    service = np.zeros(len(docs))
    results = []
    for i, doc in enumerate(docs):
        t = time.perf_counter()
        slot = server.submit(doc, id_base + i)
        if slot is None:
            raise RuntimeError("serial baseline: request does not fit the server")
        server.flush()
        while True:
            got = server.poll()
            if got:
                break
            _stalled(t, stall_s)
        z, _ = server.result(got[0], len(doc))
        service[i] = time.perf_counter() - t
        results.append(z)
    return {"service_s": service, "results": results}


def percentiles(x, qs=(50, 99)):
    return {f"p{q}": float(np.percentile(x, q)) for q in qs}
