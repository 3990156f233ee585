"""Builds ``libvfg.so`` (the C-ABI library with the gfx950 HIP kernels) in-tree.

    python -m vpfhf_amd.build            # or vpfhf_amd.build.build()

Every HIP source is compiled for gfx950 only.  Packed FP32 instructions are
disabled (``-packed-fp32-ops``) because some of their forms have produced
silent wrong results on the target machines.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libvfg.so")
SOURCES = ["vfg_u.hip", "vfg_v.hip", "vfg_misc.hip", "vfg_api.hip", "vfg_solve.hip",
           "vfg_batch.hip", "vfg_engine.hip", "vfg_mg.hip", "vfg_bmg.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "-Wno-unused-result"]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), p.stderr[-6000:]))
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed:\n%s" % p.stderr[-4000:])
    os.replace(tmp, LIB)
    # C host layer (plain C, gcc) over the C-ABI
    host = os.path.join(HERE, "libvfhost.so")
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-o", host + ".tmp",
           os.path.join(HERE, "host", "vfhost.c"), os.path.join(HERE, "host", "vfio.c"), "-L" + HERE, "-lvfg",
           "-lm", "-Wl,-rpath,$ORIGIN"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("gcc failed for vfhost.c:\n%s" % p.stderr[-4000:])
    os.replace(host + ".tmp", host)
    if verbose:
        print("built", LIB, host)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
