"""phi recovered from a u16 phi~ count and the topic's reciprocal (phi_from_count in
gpupolyaurnlda_amd/csrc/lda_core.h, the sweep's count rows) equals the f32 division phi =
(f32)x / (f32)S_k of the Phi step (polyaurn.cu:116-122) for every 16-bit count significand against
every 24-bit divisor significand -- an exhaustive check that pins the identity on the CPU; the GPU
check compares the recovered rows with the f32 rows cell by cell (gpulda_check_count_rows)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_phi_from_count_is_correctly_rounded_division(tmp_path):
    exe = tmp_path / "recip_div_check"
    subprocess.run(["g++", "-std=c++17", "-O3", "-march=native", "-ffp-contract=off", "-fopenmp", "-Wall",
                    f"-I{ROOT}/gpupolyaurnlda_amd/csrc", "-o", str(exe),
                    os.path.join(ROOT, "tests", "recip_div_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "recip div ok" in r.stdout
