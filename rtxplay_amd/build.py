"""Builds librtx.so (HIP kernels + C-ABI) in-tree for gfx950 with hipcc.

    python -m rtxplay_amd.build        (or __graft_entry__.build())

Every .hip file is compiled for gfx950 with packed-FP32 instructions disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops); the host-only C++ files
are compiled with -ffp-contract=off so that host float code rounds once per
operation, as written.  Objects go to rtxplay_amd/_build/, the library to
rtxplay_amd/librtx.so.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "librtx.so")
APP = os.path.join(PKG, "rtwo")
ARCH = os.environ.get("RTX_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["bvh_build.hip", "tlas.hip", "render.hip", "query.hip"]
CXX_SOURCES = ["capi.cpp", "objio.cpp"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
# -ffp-contract=off: every kernel rounds each operation as written, so two
# kernels that run the same arithmetic (a frame rendered alone or inside a
# batch) produce the same bits whatever the compiler's scheduling, and the
# oracle (built the same way) follows the same roundings
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH,
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-ffp-contract=off"]
HOST = ["-ffp-contract=off", "-D__HIP_PLATFORM_AMD__"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + p.stdout + p.stderr)
    return p.stderr


def build(verbose=False):
    """Compiles every source and links librtx.so; returns the library path."""
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append((obj, [cc] + COMMON + DEVICE + ["-c", os.path.join(CSRC, src), "-o", obj]))
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append((obj, [cc] + COMMON + HOST + DEVICE[:2] + ["--offload-arch=" + ARCH] +
                     ["-c", os.path.join(CSRC, src), "-o", obj]))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        logs = list(ex.map(lambda j: _run(j[1]), jobs))
    if verbose:
        for lg in logs:
            if lg.strip():
                print(lg, file=sys.stderr)
    tmp = LIB + ".tmp"
    _run([cc, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + [j[0] for j in jobs])
    os.replace(tmp, LIB)
    # the C++ batch driver (host only, plain g++ against the C-ABI)
    _run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-I" + os.path.join(ROOT, "include"),
          os.path.join(PKG, "apps", "rtwo.cpp"), "-L" + PKG, "-lrtx", "-Wl,-rpath,$ORIGIN", "-o", APP])
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
