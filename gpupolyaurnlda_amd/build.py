"""Build recipe for the native parts of gpupolyaurnlda_amd.

Produces, in-tree (so the files travel with the source tree to the GPU host):

* ``gpupolyaurnlda_amd/libGPUPolyaUrnLDA.so`` -- the drop-in HIP library for gfx950
  (the reference builds ``target/cuda/libGPUPolyaUrnLDA.so`` with nvcc, Makefile:29-30, build.rs:1-8);
* ``gpupolyaurnlda_amd/libgpulda_host.so`` -- host-only helpers (synthetic corpus generator,
  OpenMP);
* ``gpupolyaurnlda_amd/gpulda_train`` -- the C++ driver that replays the reference's train.rs call
  sequence against the library.

``python -m gpupolyaurnlda_amd.build`` builds everything; ``build_all()`` is what
``__graft_entry__.build()`` calls.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")

LIB = os.path.join(PKG, "libGPUPolyaUrnLDA.so")
HOSTLIB = os.path.join(PKG, "libgpulda_host.so")
DRIVER = os.path.join(PKG, "gpulda_train")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
OFFLOAD_ARCH = os.environ.get("GPULDA_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off on both host and device: every floating-point decision of the sampler must be
# made with the same rounded operations everywhere.  Packed FP32 instructions are disabled.
HIP_FLAGS = [
    f"--offload-arch={OFFLOAD_ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]

HIP_SOURCES = ["kernels.hip", "engine.hip"]
HOST_SOURCES = ["corpus_gen.cpp"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature name; that warning is expected.
    out = "\n".join(
        line for line in proc.stdout.splitlines() if "packed-fp32-ops" not in line
    )
    if proc.returncode != 0:
        raise RuntimeError(f"build failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def build_hip_lib() -> str:
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


def build_host_lib() -> str:
    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES]
    tmp = HOSTLIB + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-Wall",
          "-Wno-unknown-pragmas",
          f"-I{INCLUDE}", "-o", tmp, *srcs])
    os.replace(tmp, HOSTLIB)
    return HOSTLIB


def build_driver() -> str:
    srcs = [os.path.join(CSRC, s) for s in ("train_driver.cpp", "corpus_io.cpp")]
    tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unknown-pragmas", "-ffp-contract=off", f"-I{INCLUDE}", "-o", tmp, *srcs,
          f"-L{PKG}", "-lGPUPolyaUrnLDA", "-Wl,-rpath,$ORIGIN", "-pthread"])
    os.replace(tmp, DRIVER)
    return DRIVER


def build_all() -> list[str]:
    out = [build_hip_lib(), build_host_lib()]
    if os.path.exists(os.path.join(CSRC, "train_driver.cpp")):
        out.append(build_driver())
    return out


if __name__ == "__main__":
    for p in build_all():
        print(p)
