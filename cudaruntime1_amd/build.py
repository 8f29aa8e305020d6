"""Build libtrajsim.so (HIP, gfx950) in-tree, and the CPU oracle.

    python cudaruntime1_amd/build.py            # library only
    python cudaruntime1_amd/build.py --oracle   # library + oracle/

Flags: -ffp-contract=off keeps f32 rounding identical to the reference's
non-contracting CPU build (no FMA fusion), and packed-fp32 instructions are
disabled for device code.
"""
import os
import subprocess
import sys
import time
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libtrajsim.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = ["set.hip", "filter.hip", "select.hip", "dp.hip", "dp_mw.hip", "reduce.hip", "engine.hip", "server.hip",
           "compat.cpp", "dbfile.hip", "noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TS_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fopenmp", "-Wno-unused-result"]
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + cmd[-1])
    return r


def build_library(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    t0 = time.time()
    jobs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OBJ, src + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC] + COMMON + DEVICE + ["-c", path, "-o", obj]
        else:
            cmd = [HIPCC] + COMMON + ["-x", "c++", "-c", path, "-o", obj]
        jobs.append((cmd, obj))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    link = [HIPCC, "-shared", "-fPIC", "-fopenmp", "--offload-arch=" + ARCH] + [o for _, o in jobs] + ["-o", OUT + ".tmp"]
    _run(link)
    os.replace(OUT + ".tmp", OUT)
    if verbose:
        print("built %s in %.1fs" % (OUT, time.time() - t0))
    return OUT


def build_oracle(verbose=True):
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    if verbose:
        print(r.stdout.strip().splitlines()[-2] if r.stdout.strip() else "oracle built")


if __name__ == "__main__":
    build_library()
    if "--oracle" in sys.argv:
        build_oracle()
