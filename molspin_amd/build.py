"""Builds libmolspin_gpu.so (the HIP backend) in-tree for gfx950.

    python -m molspin_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmolspin_gpu.so")
ARCH = os.environ.get("MSG_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-parameter", "-Wno-unused-variable"]
# host translation units go through hipcc too (HIP runtime API); they carry no kernels
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-Wall", "-Wextra", "-Wno-unused-parameter"]

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
LIBS = ["-L" + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas", "-Wl,-rpath," + os.path.join(ROCM, "lib")]

SOURCES = [("msg_kernels.hip", "hip"), ("msg_noalloc.hip", "hip"), ("msg_abi.cpp", "host"), ("msg_resolve.cpp", "host"),
           ("msg_plan.cpp", "host"), ("msg_msd.cpp", "host")]


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src, kind in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, src + ".o")
        if kind == "hip":
            cmd = ["hipcc", "-c", "-x", "hip", path, "-o", obj] + HIP_FLAGS
        else:
            # host-only translation units: the HIP runtime API through hipcc's host compiler
            cmd = ["hipcc", "-c", path, "-o", obj] + HOST_FLAGS
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    # rocSOLVER (Hermitian eigensolver of the symmetric-decay entry) and its rocBLAS handle
    cmd = ["hipcc", "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + LIBS
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
