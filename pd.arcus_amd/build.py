"""Builds the product library libpdgpu.so for gfx950 with hipcc.

    python pd.arcus_amd/build.py

Output: pd.arcus_amd/lib/libpdgpu.so (in-tree, so it travels with the
repository to the GPU machine).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libpdgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["pdgpu.hip", "pdgpu_setup.cpp", "pdgpu_prep.cpp"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    # packed FP32 instructions are kept out of all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libpdgpu.so")
    os.replace(tmp, LIB)
    build_host_demo(verbose)
    return LIB


def build_host_demo(verbose=False):
    """C++ drop-in adapter (host/ff_gb_hip.cpp) and its driver, linked
    against libpdgpu.so with an rpath to the in-tree lib directory."""
    out = os.path.join(LIBDIR, "ff_gb_hip_demo")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(HERE, "host", "ff_gb_hip.cpp"),
           os.path.join(HERE, "host", "ff_gb_hip_demo.cpp"), "-L" + LIBDIR, "-lpdgpu",
           "-Wl,-rpath,$ORIGIN", "-o", out]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building the C++ adapter")
    return out


if __name__ == "__main__":
    print(build(verbose=True))
