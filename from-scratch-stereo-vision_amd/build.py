"""Build libstereo_mi.so (all HIP kernels + the C ABI) for gfx950 with hipcc.

    python from-scratch-stereo-vision_amd/build.py

Each csrc/*.hip file is compiled to an object in csrc/_build/ and the
objects are linked into one shared library next to this file, so the built
library travels with the source tree.  Files whose results must be bit-exact
against float64 references are compiled with -ffp-contract=off.
"""

import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libstereo_mi.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
          # packed FP32 instructions are kept out of every kernel
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]
# per-file floating-point contraction: exact files never fuse multiply-adds
EXACT = {"refine.hip", "common.hip", "planes.hip", "imageio.hip", "evalmetrics.hip",
         "geometry.hip", "mesh.hip", "splat.hip"}


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    t0 = time.time()
    for src in sources():
        obj = os.path.join(OBJ, src[:-4] + ".o")
        contract = "-ffp-contract=off" if src in EXACT else "-ffp-contract=fast"
        cmd = [HIPCC] + COMMON + [contract, "-I", CSRC, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((src, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    objs = []
    failed = []
    for src, obj, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
        objs.append(obj)
    if failed:
        for src, out in failed:
            sys.stderr.write("=== %s ===\n%s\n" % (src, out))
        raise RuntimeError("hipcc failed for: " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    if verbose:
        print("built %s in %.1f s" % (LIB, time.time() - t0))
    return LIB


if __name__ == "__main__":
    build(verbose=True)
