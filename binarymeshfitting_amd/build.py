"""Builds libbmf.so (HIP, gfx950) in-tree with hipcc.

Flags: -ffp-contract=off keeps every multiply and add separately rounded
(explicit __builtin_fmaf where a fused op is wanted), correctly rounded FP32
divide/sqrt and denormals kept, packed FP32 disabled.
"""
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libbmf.so"
SOURCES = ["bmf_kernels.hip", "bmf_capi.hip", "bmf_diff.hip", "bmf_dmc.hip", "bmf_format.hip", "bmf_lod.hip", "bmf_stitch.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def build(verbose=False, extra=()):
    objs = []
    build_dir = PKG / "_build"
    build_dir.mkdir(exist_ok=True)
    for src in SOURCES:
        obj = build_dir / (src + ".o")
        cmd = [HIPCC, *FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(str(obj))
    tmp = str(LIB) + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


HOST = PKG / "host"
HOST_LIB = PKG / "libbmf_host.so"
HOST_DEMO = HOST / "bmf_host_demo"


def build_host(verbose=False):
    """The C++ host layer (g++) over the C ABI, and its demo driver."""
    rocm_lib = "/opt/rocm/lib"
    common = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-fPIC", f"-I{PKG.parent / 'include'}"]
    link = [f"-L{PKG}", "-lbmf", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath-link,{rocm_lib}",
            f"-L{rocm_lib}"]
    cmds = [
        [*common, "-shared", "-o", str(HOST_LIB), str(HOST / "bmf_host.cpp"), *link],
        [*common, "-o", str(HOST_DEMO), str(HOST / "bmf_host_demo.cpp"), str(HOST / "bmf_host.cpp"), *link,
         "-Wl,-rpath,$ORIGIN/.."],
    ]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
