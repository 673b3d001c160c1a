"""Writes tests/golden/philox_rocrand.json: Philox4x32-10 blocks from rocRAND's open-source
engine header (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h), compiled for the host."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

with tempfile.TemporaryDirectory() as td:
    exe = os.path.join(td, "philox_rocrand")
    subprocess.run(["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    os.path.join(HERE, "philox_rocrand.cpp"), "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
    f.write(out)
print("wrote philox_rocrand.json")
