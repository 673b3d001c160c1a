"""Build recipe of the CPU oracle (test infrastructure): oracle/liblda_oracle.so.

The reference's own sources cannot be built here (CUDA 8 + cuRAND/cuBLAS, sm_61, and a Rust host
that needs cargo), so there is no oracle/_ref build; the oracle is the C restatement in
lda_oracle.c.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liblda_oracle.so")


def build() -> str:
    tmp = SO + ".tmp"
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-Wall",
                    "-o", tmp, os.path.join(HERE, "lda_oracle.c"), "-lm"], check=True)
    os.replace(tmp, SO)
    return SO


if __name__ == "__main__":
    print(build())
