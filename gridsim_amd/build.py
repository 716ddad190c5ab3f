"""Builds libgridsim_hip.so in-tree with hipcc for gfx950 (the engine) and the CPU oracle (tests only).

    python -m gridsim_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libgridsim_hip.so")
ARCH = os.environ.get("GRIDSIM_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fvisibility=hidden", "-Wall", "-Wno-unused-function",
          "-I", INCLUDE, "-I", CSRC]
# -packed-fp32-ops: no v_pk_*_f32 (the host pass warns that it does not know the feature; harmless)
HIP_ONLY = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-fno-gpu-rdc"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_engine(verbose: bool = True) -> str:
    out_dir = tempfile.mkdtemp(prefix="gship_", dir=os.environ.get("TMPDIR", "/tmp"))
    objs = []
    jobs = []
    for src, kind in (("model.cpp", "host"), ("engine.cpp", "hip"), ("server.cpp", "hip"), ("nr_kernels.hip", "hip")):
        obj = os.path.join(out_dir, src + ".o")
        path = os.path.join(CSRC, src)
        if kind == "host":
            cmd = ["g++"] + COMMON + ["-c", path, "-o", obj]
        else:
            cmd = [hipcc(), "-x", "hip"] + COMMON + HIP_ONLY + ["-c", path, "-o", obj]
        jobs.append(cmd)
        objs.append(obj)
    procs = []
    for cmd in jobs:
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    rc = [p.wait() for p in procs]
    if any(rc):
        raise RuntimeError(f"compilation failed: {rc}")
    tmp_lib = LIB + ".tmp"
    _run([hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp_lib] + objs + ["-lpthread"])
    os.replace(tmp_lib, LIB)
    shutil.rmtree(out_dir, ignore_errors=True)
    return LIB


def build_oracle() -> None:
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s"], check=True)


def main(argv=None):
    build_engine()
    build_oracle()
    print("built", LIB)


if __name__ == "__main__":
    main(sys.argv[1:])
