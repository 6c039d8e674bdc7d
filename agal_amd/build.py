"""Builds the HIP shared library in-tree (agal_amd/lib/libagal.so) with hipcc for gfx950."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
SOURCES = ["capi.hip", "lbm_kernels.hip", "amr_kernels.hip", "geometry.hip", "links.hip", "predictor.hip"]

# -ffp-contract=off: no FMA contraction, so FP32/FP64 results are reproducible operation for
# operation (the refinement flags depend on them).  The packed-FP32 feature is disabled for all
# device code.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
]


def hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.sep not in cand or os.path.exists(cand):
            return cand
    return "hipcc"


def build_library(verbose: bool = False, jobs: int = 0) -> str:
    """Compiles every source (one hipcc process per source, in parallel), then links the objects."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    out = os.path.join(LIBDIR, "libagal.so")
    tmp = out + ".tmp"
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    cmds = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmds.append(([hipcc()] + compile_flags + ["-c", "-o", obj, os.path.join(CSRC, src)], obj))
    jobs = jobs or min(len(cmds), max(1, min(16, os.cpu_count() or 1)))
    procs, objs = [], []
    pending = list(cmds)
    while pending or procs:
        while pending and len(procs) < jobs:
            cmd, obj = pending.pop(0)
            if verbose:
                print(" ".join(cmd))
            procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True), cmd))
            objs.append(obj)
        p, cmd = procs.pop(0)
        txt, _ = p.communicate()
        if p.returncode != 0:
            for q, _c in procs:
                q.wait()
            raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + txt)
    link = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc link failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
