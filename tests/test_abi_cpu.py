"""CPU-side checks of the boundary: the library loads, exports exactly what include/*.h declares,
the struct layouts match the reference's repr(C) structs (src/ffi.rs:4-13, src/buffer.rs:4-19),
the train.rs buffer packing is restated faithfully, and the corpus generator is deterministic.
No GPU call is made here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for h in ("gpulda_abi.h", "gpulda_ext.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**([a-z_][a-z0-9_]*)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    from gpupolyaurnlda_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "build the library first"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    declared = _declared_functions()
    assert {"initialize", "sample_phi", "sample_z_async", "sync_buffer", "cleanup"} <= declared
    missing = declared - exported
    assert not missing, missing
    # the ctypes table covers exactly the declared set
    assert set(_lib.EXPORTED) == declared


def test_library_loads_without_gpu():
    from gpupolyaurnlda_amd import _lib
    lib = _lib.lib()
    assert lib.gpulda_set_seed(5) == 0 and lib.gpulda_get_seed() == 5
    assert lib.gpulda_set_seed(0) == 0


def test_struct_layouts_match_reference():
    from gpupolyaurnlda_amd._lib import Args, Buffer
    assert C.sizeof(Args) == 32
    assert [Args.alpha.offset, Args.beta.offset, Args.K.offset, Args.V.offset, Args.C.offset,
            Args.buffer_size.offset, Args.max_D.offset] == [0, 4, 8, 12, 16, 24, 28]
    assert C.sizeof(Buffer) == 96
    offs = [getattr(Buffer, f).offset for f, _ in Buffer._fields_]
    assert offs == [0, 8, 16, 24, 32, 36, 40, 48, 56, 64, 72, 80, 88]


def test_header_static_asserts_compile():
    src = '#include "gpulda_abi.h"\n#include "gpulda_ext.h"\nint main(){return 0;}\n'
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{ROOT}/include", "-x", "c++", "-"], input=src,
                   text=True, check=True)
    subprocess.run(["gcc", "-std=c11", "-fsyntax-only", f"-I{ROOT}/include", "-x", "c", "-"], input=src,
                   text=True, check=True)


def test_fill_buffer_restates_train_rs():
    from gpupolyaurnlda_amd.train import fill_buffer
    d = np.array([3, 4, 2, 5, 1, 1], np.uint32)
    # stops before exceeding buffer_size (train.rs:30) or max_D
    assert fill_buffer(d, 0, buffer_size=9, max_D=100) == (3, 9)
    assert fill_buffer(d, 3, buffer_size=9, max_D=100) == (3, 7)
    assert fill_buffer(d, 0, buffer_size=100, max_D=2) == (2, 7)
    assert fill_buffer(d, 6, buffer_size=100, max_D=2) == (0, 0)
    # train.rs:27-29 panics as soon as it reads a document larger than a buffer
    with pytest.raises(ValueError):
        fill_buffer(np.array([3, 4, 2, 10], np.uint32), 0, buffer_size=9, max_D=100)
    # empty documents pass through instead of ending the iteration
    assert fill_buffer(np.array([2, 0, 3], np.uint32), 0, 100, 100) == (3, 5)


def _fill_buffer_loop(d, doc_pos, buffer_size, max_D):
    # train.rs:19-45 document by document: check the size, then stop when it would not fit
    n_docs = n_tokens = 0
    while doc_pos + n_docs < len(d):
        size = int(d[doc_pos + n_docs])
        if size > buffer_size:
            raise ValueError(size)
        if n_docs + 1 > max_D or n_tokens + size > buffer_size:
            break
        n_docs += 1
        n_tokens += size
    return n_docs, n_tokens


def test_fill_buffer_matches_document_loop():
    from gpupolyaurnlda_amd.train import fill_buffer
    rng = np.random.default_rng(7)
    for _ in range(2000):
        D = int(rng.integers(0, 30))
        d = rng.integers(0, 20, D).astype(np.uint32)
        args = (int(rng.integers(0, D + 1)), int(rng.integers(0, 40)), int(rng.integers(1, 10)))
        try:
            want = _fill_buffer_loop(d, *args)
        except ValueError:
            with pytest.raises(ValueError):
                fill_buffer(d, *args)
            continue
        assert fill_buffer(d, *args) == want, (d, args)


def test_generator_is_deterministic_and_shaped():
    from gpupolyaurnlda_amd import corpus
    spec = corpus.TINY
    a = corpus.generate(spec)
    b = corpus.generate(spec)
    assert np.array_equal(a["w"], b["w"]) and np.array_equal(a["z"], b["z"])
    assert len(a["d_len"]) == spec.D and a["d_len"].min() >= 1
    assert abs(a["d_len"].mean() - 50) < 1.5
    assert a["w"].max() < spec.V and a["z"].max() < spec.K
    med = corpus.lengths(corpus.MEDIUM)
    assert med.max() <= 2000 and abs(np.median(med) - 100) <= 3
    bat = corpus.lengths(corpus.BATCHED.scaled(200000))
    assert bat.min() >= 4 and bat.max() <= 20000


def test_slice_hash_sees_every_word():
    """The slice cache's default mode reuses a device copy only when the whole content matches: a
    refilled host buffer that differs in one word -- one the old sampled fingerprint never read --
    or in one document length must hash differently; equal content at another address hashes
    the same.  (The reference re-uploads every call, train.cu:160-163.)"""
    import ctypes as C
    import time
    from gpupolyaurnlda_amd._lib import lib, ptr
    rng = np.random.default_rng(0)
    n, D = 3_000_000, 30_000
    w = rng.integers(0, 50_000, n).astype(np.uint32)
    d = np.full(D, n // D, np.uint32)

    def h(w, d):
        return lib().gpulda_slice_hash(ptr(w, C.c_uint32), ptr(d, C.c_uint32), len(d), len(w))

    base = h(w, d)
    assert h(w.copy(), d.copy()) == base
    step = n // 4096
    for pos in (1, step + 1, n // 2 + 3, n - 2):  # none of them on the sampled grid
        w2 = w.copy()
        w2[pos] ^= 1
        assert h(w2, d) != base, pos
    d2 = d.copy()
    d2[5] += 1
    d2[6] -= 1
    assert h(w, d2) != base
    assert h(w[:-1], d) != base
    big = rng.integers(0, 100_000, 50_000_000).astype(np.uint32)
    t0 = time.perf_counter()
    h(big, np.array([len(big)], np.uint32))
    print(f"slice hash: {len(big) * 4 / (time.perf_counter() - t0) / 1e9:.1f} GB/s")


def test_stage_map_and_pick_reject_tables_the_device_state_cannot_hold():
    """The doc-topic state on the device is one count per topic plus a table of the topics present
    (at most cap - 1 of them, 2 * topic per 16-bit slot): the stage entry points refuse inputs outside
    that before any device work, with a message."""
    from gpupolyaurnlda_amd._lib import lib
    L = lib()

    def err():
        return L.gpulda_last_error().decode()

    def stage_map(ops, cap, K):
        ops = np.ascontiguousarray(ops, dtype=np.uint32)
        out = np.zeros(cap, np.uint32)
        return L.gpulda_stage_map(ops.ctypes.data_as(C.POINTER(C.c_uint32)), len(ops), cap, K,
                                  out.ctypes.data_as(C.POINTER(C.c_uint32)))

    assert stage_map(np.arange(64), 64, 1024) != 0 and "no free slot" in err()
    assert stage_map(np.zeros(65536), 64, 1024) != 0 and "past 65535" in err()
    assert stage_map([1, 2], 64, 32768) != 0 and "out of range" in err()
    assert stage_map([1, 5000], 64, 1024) != 0 and ">= K" in err()

    table = np.zeros(64, np.uint32)
    table[3] = (7 << 16) | 2
    table[9] = (7 << 16) | 1
    row = np.ones(1024, np.float32)
    u = np.full(1, 0.5, np.float32)
    out = np.zeros(1, np.uint32)
    rc = L.gpulda_stage_pick(table.ctypes.data_as(C.POINTER(C.c_uint32)), 64, 1024,
                             row.ctypes.data_as(C.POINTER(C.c_float)), C.c_float(1.0),
                             u.ctypes.data_as(C.POINTER(C.c_float)), u.ctypes.data_as(C.POINTER(C.c_float)), 1, 1,
                             out.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert rc != 0 and "repeats topic 7" in err()
