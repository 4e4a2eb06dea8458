"""Streaming-copy yardstick for the hardware byte counters and the achievable HBM rate.

    python py-orbit_amd/calibrate_copy.py [--mib 1024] [--reps 20]

Copies a buffer of the given size (read + write) `reps` times with a plain device
kernel and prints the achieved bandwidth; --wide also sweeps copies with 16-byte loads,
several in flight per thread, plain or non-temporal stores, and prints the best rate.  Under `rocprofv3 --pmc FETCH_SIZE` (or
WRITE_SIZE) the per-dispatch counter of `copy_kernel` should read the buffer size; the
ratio of the two is the scale factor applied to the other kernels' counters.
"""
import argparse
import ctypes as C
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _obnative as N  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mib", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--wide", action="store_true",
                    help="also sweep the 16-byte multi-load copy variants and print the best")
    a = ap.parse_args()
    n = a.mib * 1024 * 1024 // 8
    ms = C.c_double()
    N.call("ob_calibrate_copy", n, a.reps, C.byref(ms))
    gbs = 2 * 8 * n / (ms.value * 1e-3) / 1e9
    print(f"copy {a.mib} MiB: {ms.value:.4f} ms per copy, {gbs:.1f} GB/s (read+write), "
          f"{8 * n // 1024} KiB read and written per dispatch")
    if a.wide:
        v = C.c_int()
        N.call("ob_calibrate_copy_wide", n, a.reps, C.byref(ms), C.byref(v))
        gbs = 2 * 8 * n / (ms.value * 1e-3) / 1e9
        grid = (1024, 2048, 4096, 8192)[v.value // 4]
        print(f"wide copy {a.mib} MiB: {ms.value:.4f} ms per copy, {gbs:.1f} GB/s (read+write); "
              f"best variant: {grid} workgroups, {8 if v.value & 1 else 4} x 16-byte loads in "
              f"flight per thread, {'non-temporal' if v.value & 2 else 'plain'} stores")


if __name__ == "__main__":
    main()
