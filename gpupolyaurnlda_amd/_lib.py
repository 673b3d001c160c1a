"""ctypes bindings of libGPUPolyaUrnLDA.so (include/gpulda_abi.h, include/gpulda_ext.h).

The library is loaded from this package directory.  There is no fallback: if the HIP library is
missing, every entry point raises ``RuntimeError`` naming the build step.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libGPUPolyaUrnLDA.so")
HOSTLIB_PATH = os.path.join(PKG, "libgpulda_host.so")

u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)


class Args(C.Structure):
    """repr(C) Args of src/ffi.rs:4-13 (== src/cu/train.cuh:13-21), 32 bytes."""

    _fields_ = [
        ("alpha", C.c_float),
        ("beta", C.c_float),
        ("K", C.c_uint32),
        ("V", C.c_uint32),
        ("C", u32p),
        ("buffer_size", C.c_uint32),
        ("max_D", C.c_uint32),
    ]


class Buffer(C.Structure):
    """repr(C) Buffer of src/buffer.rs:4-19 (== src/cu/train.cuh:23-37), 96 bytes."""

    _fields_ = [
        ("z", u32p),
        ("w", u32p),
        ("d", u32p),
        ("K_d", u32p),
        ("n_docs", C.c_uint32),
        ("n_tokens", C.c_uint32),
        ("gpu_z", C.c_void_p),
        ("gpu_w", C.c_void_p),
        ("gpu_d_len", C.c_void_p),
        ("gpu_d_idx", C.c_void_p),
        ("gpu_K_d", C.c_void_p),
        ("gpu_rng", C.c_void_p),
        ("stream", C.c_void_p),
    ]


assert C.sizeof(Args) == 32 and Args.buffer_size.offset == 24 and Args.max_D.offset == 28
assert C.sizeof(Buffer) == 96 and Buffer.n_docs.offset == 32 and Buffer.stream.offset == 88


class GenSpec(C.Structure):
    _fields_ = [
        ("V", C.c_uint32),
        ("K", C.c_uint32),
        ("D", C.c_uint32),
        ("len_kind", C.c_uint32),
        ("len_a", C.c_double),
        ("len_b", C.c_double),
        ("len_c", C.c_double),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("zipf_s", C.c_double),
        ("seed", C.c_uint64),
    ]


_SIGS = {
    # five-call ABI
    "initialize": (None, [C.POINTER(Args), C.POINTER(Buffer), C.c_uint32]),
    "sample_phi": (None, []),
    "sample_z_async": (None, [C.POINTER(Buffer)]),
    "sync_buffer": (None, [C.POINTER(Buffer)]),
    "cleanup": (None, [C.POINTER(Buffer), C.c_uint32]),
    # extensions
    "gpulda_set_seed": (C.c_int, [C.c_uint64]),
    "gpulda_get_seed": (C.c_uint64, []),
    "gpulda_last_error": (C.c_char_p, []),
    "gpulda_device_info": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int), u64p, C.POINTER(C.c_int)]),
    "gpulda_serve_set_deadline": (C.c_int, [C.c_double]),
    "gpulda_set_slice_cache": (C.c_int, [C.c_int]),
    "gpulda_set_slice_cache_budget": (C.c_int, [C.c_uint64]),
    "gpulda_slice_cache_stats": (C.c_int, [u32p, u64p, u64p]),
    "gpulda_slice_hash": (C.c_uint64, [u32p, u32p, C.c_uint32, C.c_uint64]),
    "gpulda_iteration": (C.c_uint32, []),
    "gpulda_copy_phi": (C.c_int, [f32p]),
    "gpulda_copy_alias": (C.c_int, [f32p, u32p]),
    "gpulda_copy_sigma_a": (C.c_int, [f32p]),
    "gpulda_copy_counts": (C.c_int, [u32p]),
    "gpulda_set_counts": (C.c_int, [u32p]),
    "gpulda_set_iteration": (C.c_int, [C.c_uint32]),
    "gpulda_set_part_tokens": (C.c_int, [C.c_uint64]),
    "gpulda_set_count_rows": (C.c_int, [C.c_int]),
    "gpulda_set_f32_rows": (C.c_int, [C.c_uint32]),
    "gpulda_check_count_rows": (C.c_int, [C.POINTER(C.c_int), u64p, u64p]),
    "gpulda_copy_topic_sums": (C.c_int, [u64p]),
    "gpulda_copy_poisson_tables": (C.c_int, [f32p, u32p]),
    "gpulda_sweep_stats": (C.c_int, [u64p]),
    "gpulda_reset_stats": (C.c_int, []),
    "gpulda_corpus_load": (C.c_int, [u32p, u32p, C.c_uint32, u32p]),
    "gpulda_corpus_counts_from_z": (C.c_int, []),
    "gpulda_corpus_copy_z": (C.c_int, [u32p]),
    "gpulda_corpus_iterate": (C.c_int, [C.c_uint32, f32p]),
    "gpulda_corpus_loglik": (C.c_int, [f64p]),
    "gpulda_corpus_schedule": (C.c_int, [C.c_int, C.c_uint32, C.c_uint32]),
    "gpulda_corpus_use_graph": (C.c_int, [C.c_int]),
    "gpulda_corpus_autotune": (C.c_int, [f32p, C.c_uint32, u32p, C.POINTER(C.c_int)]),
    "gpulda_corpus_release": (C.c_int, []),
    "gpulda_calibrate_copy": (C.c_int, [C.c_uint64, C.c_uint32, f32p]),
    "gpulda_stage_philox": (C.c_int, [u32p, u32p, u32p, C.c_uint32]),
    "gpulda_stage_scan64": (C.c_int, [f32p, f32p, C.c_uint32]),
    "gpulda_stage_map": (C.c_int, [u32p, C.c_uint32, C.c_uint32, C.c_uint32, u32p]),
    "gpulda_stage_alias": (C.c_int, [f32p, C.c_uint32, C.c_uint32, u64p, C.c_float, f32p, f32p, f32p, u32p]),
    "gpulda_stage_alias_draw": (C.c_int, [f32p, u32p, C.c_uint32, u32p, u32p, C.c_uint32, u32p]),
    "gpulda_stage_pick": (C.c_int, [u32p, C.c_uint32, C.c_uint32, f32p, C.c_float, f32p, f32p, C.c_uint32, C.c_int,
                                    u32p]),
    "gpulda_stage_doc_offsets": (C.c_int, [u32p, C.c_uint32, u64p]),
    "gpulda_stage_poisson": (C.c_int, [u32p, C.c_uint32, C.c_float, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, u32p]),
    "gpulda_stage_phi_draw": (C.c_int, [u32p, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32, C.c_uint64, f32p, u64p]),
    "gpulda_stage_sweep": (C.c_int, [u32p, u32p, u64p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                                     C.c_uint32, f32p, f32p, f32p, u32p, u64p]),
    "gpulda_stage_count": (C.c_int, [u32p, u32p, C.c_uint64, C.c_uint32, C.c_uint32, u32p]),
    "gpulda_stage_foldin": (C.c_int, [u32p, u64p, C.c_uint32, u32p, C.c_uint32, C.c_uint64, C.c_uint32, C.c_uint32,
                                      f32p, f32p, f32p, u32p, u32p, u64p]),
    "gpulda_foldin_noalloc": (C.c_int, [u32p, u64p, C.c_uint32, u32p, C.c_uint32, u32p, u64p]),
    "gpulda_serve_open": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
    "gpulda_serve_submit": (C.c_int, [u32p, C.c_uint32, C.c_uint32, u32p]),
    "gpulda_serve_flush": (C.c_int, []),
    "gpulda_serve_poll": (C.c_int, [u32p, C.c_uint32, u32p]),
    "gpulda_serve_result": (C.c_int, [C.c_uint32, u32p, u32p]),
    "gpulda_serve_stats": (C.c_int, [u32p, u32p, u64p]),
    "gpulda_serve_close": (C.c_int, []),
}

# Declared in include/*.h; the "library exports every declared symbol" test reads this list.
EXPORTED = sorted(_SIGS)

_HOST_SIGS = {
    "gpulda_gen_lengths": (C.c_uint64, [C.POINTER(GenSpec), u32p]),
    "gpulda_gen_tokens": (C.c_int, [C.POINTER(GenSpec), u32p, u32p, u32p, f32p]),
    "gpulda_gen_requests": (C.c_uint64, [f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                         C.c_uint32, C.c_uint32, C.c_double, C.c_uint64, u32p, f64p, u32p]),
}

_lib = None
_host = None


def _load(path, sigs, what):
    if not os.path.exists(path):
        raise RuntimeError(
            f"{what} not found at {path}; build it with `python -m gpupolyaurnlda_amd.build` "
            "(or __graft_entry__.build())"
        )
    lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
    for name, (res, argt) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = argt
    return lib


def lib():
    """The HIP library (loaded once)."""
    global _lib
    if _lib is None:
        _lib = _load(LIB_PATH, _SIGS, "libGPUPolyaUrnLDA.so")
    return _lib


def hostlib():
    """Host helpers (corpus generator)."""
    global _host
    if _host is None:
        _host = _load(HOSTLIB_PATH, _HOST_SIGS, "libgpulda_host.so")
    return _host


def ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays passed to the library must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().gpulda_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed: {msg}")
