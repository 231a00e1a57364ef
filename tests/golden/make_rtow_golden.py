"""Writes tests/golden/rtow_ref.npz from the reference CPU tracer.

Runs oracle/_ref/rtow (built from the reference's rtow.cxx by oracle/ref.mk,
in the container that holds the reference tree) and stores, from its PPM
output (1280x720, 10 spp, depth 50, FP64, glibc rand() without srand):
  sha256 of the w*h*3 pixel bytes (PPM order, top row first),
  every 45th row of pixels, and the per-channel byte sums.
Usage:  make -f oracle/ref.mk && python tests/golden/make_rtow_golden.py
"""
import hashlib
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "rtow")
    if not os.path.exists(exe):
        sys.exit("build the reference first: make -f oracle/ref.mk")
    txt = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    assert txt[0] == "P3"
    w, h, maxval = int(txt[1]), int(txt[2]), int(txt[3])
    px = np.array(txt[4:4 + w * h * 3], dtype=np.int64)
    assert maxval == 255 and px.size == w * h * 3 and px.max() <= 255
    img = px.astype(np.uint8).reshape(h, w, 3)
    rows = np.arange(0, h, 45)
    np.savez_compressed(
        os.path.join(ROOT, "tests", "golden", "rtow_ref.npz"),
        w=w, h=h, spp=10, depth=50,
        sha256=np.frombuffer(hashlib.sha256(img.tobytes()).digest(), np.uint8),
        rows=rows, row_pixels=img[rows],
        channel_sums=img.reshape(-1, 3).sum(axis=0, dtype=np.int64),
    )
    print("wrote rtow_ref.npz", w, h, hashlib.sha256(img.tobytes()).hexdigest())


if __name__ == "__main__":
    main()
