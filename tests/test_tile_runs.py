"""Run tables of the n_wk reduction's tile reorders (gpupolyaurnlda_amd/csrc/tile_runs.h) on the
CPU: the place the reorder kernel computes for each rank (bitmap word prefix, popcount, run table)
equals the place walking the output gave it, for identity, random and word-sorted output orders of
sizes around the 32,768-entry tile; and the parallel per-tile build of both passes gives the ranks
and places of a sequential walk of the outputs.  The header is plain C++; the check program is compiled with the
host compiler."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_tile_run_places(tmp_path):
    exe = tmp_path / "tile_runs_check"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-pthread", f"-I{ROOT}/gpupolyaurnlda_amd/csrc", "-o", str(exe),
                    os.path.join(ROOT, "tests", "tile_runs_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "tile runs ok" in r.stdout
