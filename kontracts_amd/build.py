"""Build the engine's shared library in-tree with hipcc for gfx950.

    python -m kontracts_amd.build        # -> kontracts_amd/libkontracts_amd.so

Every HIP translation unit is compiled with packed FP32 ops disabled; host-only
C++ units are compiled by hipcc as plain C++ against the HIP headers.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libkontracts_amd.so")
ARCH = os.environ.get("KT_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the engine needs ROCm's hipcc to build")


def _sources():
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    return hip, cpp


def _compile(args):
    cmd, src = args
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = ("packed-fp32-ops",)
    err = "\n".join(l for l in res.stderr.splitlines() if not any(n in l for n in noise))
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{err}")
    return src, err


def build(verbose: bool = False) -> str:
    hipcc = _hipcc()
    hip, cpp = _sources()
    objdir = tempfile.mkdtemp(prefix="kt_build_")
    try:
        jobs = []
        objs = []
        for f in hip:
            o = os.path.join(objdir, f + ".o")
            objs.append(o)
            jobs.append(([hipcc, *HIP_FLAGS, "-I", INCLUDE, "-c", os.path.join(CSRC, f), "-o", o], f))
        for f in cpp:
            o = os.path.join(objdir, f + ".o")
            objs.append(o)
            jobs.append(([hipcc, *HOST_FLAGS, "-I", INCLUDE, "-c", os.path.join(CSRC, f), "-o", o], f))
        workers = min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)), 16)
        with cf.ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
            for src, err in ex.map(_compile, jobs):
                if verbose and err.strip():
                    print(f"[{src}]\n{err}", file=sys.stderr)
        tmp = LIB + ".tmp"
        link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
        res = subprocess.run(link, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed:\n{' '.join(link)}\n{res.stderr}")
        os.replace(tmp, LIB)
    finally:
        shutil.rmtree(objdir, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
