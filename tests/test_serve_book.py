"""Fold-in server bookkeeping (gpupolyaurnlda_amd/csrc/serve_book.h) on the CPU: finished and
lost reports, the per-call limit, quarantine until collected and no worker runs, and reuse of a
released slot.  The header is plain C++; the check program is compiled with the host compiler."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_serve_bookkeeping(tmp_path):
    exe = tmp_path / "serve_book_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", f"-I{ROOT}/gpupolyaurnlda_amd/csrc", "-o", str(exe),
                    os.path.join(ROOT, "tests", "serve_book_check.cpp")], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "serve book ok" in r.stdout
