"""Builds libxenon_hip.so for gfx950 in-tree (xenon_amd/lib/).

Device code: hipcc --offload-arch=gfx950 with -ffp-contract=off (no fused
multiply-add anywhere, so +,-,* round as the reference's host code does;
HIP's default correctly rounded fp32 division and square root are kept) and
packed-FP32 instructions disabled. Host code (scene builder and C ABI): g++
with -ffp-contract=off against the HIP runtime headers.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "lib", "libxenon_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

# -amdgpu-atomic-optimizer-strategy=None: every global atomic here is
# already issued by one lane with the wave's total; the optimizer's rewrite
# of such an atomic reads its result right after issuing it, which
# serialised the shading kernel's two appends of a pass.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-amdgpu-atomic-optimizer-strategy=None",
             "-Wno-unused-command-line-argument"]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-pthread", "-D__HIP_PLATFORM_AMD__",
             f"-I{ROCM}/include"]

DEVICE_SOURCES = ["device/kernels.hip", "device/guide.hip", "device/image.hip", "device/bvh.hip"]
HOST_SOURCES = ["host/scene.cpp", "host/scene_api.cpp", "host/scene_io.cpp", "host/api.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compiles every source and links the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    inc = [f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
    procs, objs = [], []
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmd = [HIPCC, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        cmd = ["g++", *CXX_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    failed = False
    for cmd, p in procs:
        out, err = p.communicate()
        if verbose:
            print(" ".join(cmd), flush=True)
        if p.returncode != 0:
            sys.stderr.write(out + err)
            failed = True
    if failed:
        raise RuntimeError("compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
