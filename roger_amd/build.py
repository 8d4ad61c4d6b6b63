"""Builds libroger_amd.so (all HIP sources, gfx950) in-tree with hipcc.

Each source compiles to an object in its own hipcc process (in parallel, the
tier-1 kernels take minutes), then one link produces the shared library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libroger_amd.so")
OBJ = os.path.join(HERE, "_obj")
SOURCES = ["capi.hip", "colour.hip", "dwt.hip", "t1_encode.hip", "t1_mq.hip", "t1_chain.hip", "t1_decode.hip",
           "t1_rate.hip", "batch.hip", "t2_write.hip", "t2_read.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",                       # FP32 op order is part of the parity contract
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
    return r.returncode


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    cmds = [[HIPCC] + FLAGS + ["-c", "-o", o, os.path.join(CSRC, s)] for s, o in zip(SOURCES, objs)]
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        rcs = list(ex.map(lambda c: _run(c, verbose), cmds))
    if any(rcs):
        raise RuntimeError("hipcc failed compiling libroger_amd sources")
    tmp = LIB + ".tmp"
    if _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose):
        raise RuntimeError("hipcc failed linking libroger_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
