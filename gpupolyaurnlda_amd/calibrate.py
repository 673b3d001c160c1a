"""Plain copy kernel of known size (``python -m gpupolyaurnlda_amd.calibrate [MiB] [reps]``).

Used to scale rocprofv3's FETCH_SIZE / WRITE_SIZE counters: one copy reads and writes exactly
MiB mebibytes, so counter / expected gives the factor applied to the sweep kernel's counts."""
import ctypes as C
import sys

from ._lib import check, lib


def main():
    mib = int(sys.argv[1]) if len(sys.argv) > 1 else 1024
    reps = int(sys.argv[2]) if len(sys.argv) > 2 else 5
    ms = C.c_float(0)
    check(lib().gpulda_calibrate_copy(mib << 20, reps, C.byref(ms)), "calibrate_copy")
    gbs = 2 * (mib << 20) / (ms.value * 1e-3) / 1e9
    print(f"copy {mib} MiB x {reps}: {ms.value:.3f} ms per copy, {gbs:.1f} GB/s (read + write)")


if __name__ == "__main__":
    main()
