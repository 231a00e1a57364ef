"""Writes tests/golden/frand48_kat.npz: Frand48 known answers from glibc.

The device RNG (optx/frand48.h:17-35) is the POSIX drand48 generator: a
48-bit LCG (a = 0x5DEECE66D, c = 0xB) whose output x/2^48 is rounded to
float.  glibc's erand48 steps the same recurrence on a caller-held 48-bit
state and returns x/2^48 as a double (exact), so float32(erand48()) is the
reference value.  init(seed) sets the state to seed and discards 16 draws.
This script uses only glibc, not this repository's code.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SEEDS = list(range(64)) + [4711, 1 << 20, 921599, 2073599, 123456789, 0xFFFFFFFF]
K = 64


def main():
    libc = ctypes.CDLL("libc.so.6")
    libc.erand48.restype = ctypes.c_double
    libc.erand48.argtypes = [ctypes.POINTER(ctypes.c_ushort * 3)]
    out = np.empty((len(SEEDS), K), np.float32)
    for r, seed in enumerate(SEEDS):
        x = (ctypes.c_ushort * 3)(seed & 0xFFFF, (seed >> 16) & 0xFFFF, 0)
        for _ in range(16):
            libc.erand48(ctypes.byref(x))
        for j in range(K):
            out[r, j] = np.float32(libc.erand48(ctypes.byref(x)))
    np.savez_compressed(os.path.join(HERE, "frand48_kat.npz"), seeds=np.array(SEEDS, np.uint32), draws=out)
    print("wrote frand48_kat.npz", out.shape)


if __name__ == "__main__":
    main()
