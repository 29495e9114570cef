"""Builds neurongpu_amd/libneurongpu.so in-tree.

HIP kernels (csrc/*.hip) are compiled by hipcc for gfx950; host code
(csrc/*.cpp) by g++ against the HIP runtime headers.  FP contraction is off
everywhere: the kernels spell out every fused multiply-add.

Usage: python -m neurongpu_amd.build   (or build() from Python)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libneurongpu.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("NGPU_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]
CXX_FLAGS = [
    "-O2", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off",
    "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"),
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC, "-Wall",
    "-Wno-unused-function",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    hip_src = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp_src = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    procs, objs = [], []
    for f in hip_src:
        o = os.path.join(OBJ, f + ".o")
        cmd = [os.path.join(ROCM, "bin", "hipcc"), *HIP_FLAGS, "-c",
               os.path.join(CSRC, f), "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), f))
        objs.append(o)
    for f in cpp_src:
        o = os.path.join(OBJ, f + ".o")
        cmd = ["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, f), "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), f))
        objs.append(o)
    failed = [f for p, f in procs if p.wait() != 0]
    if failed:
        raise RuntimeError("compilation failed: " + ", ".join(failed))
    tmp = LIB + ".tmp"
    _run([os.path.join(ROCM, "bin", "hipcc"), "--offload-arch=" + ARCH,
          "-shared", "-fPIC", "-fopenmp", "-o", tmp, *objs,
          "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")], verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
