"""Builds libpsr_dcfr.so (host game build + gfx950 DCFR kernels + C ABI) in
place with hipcc.  Usage: python -m poker_solver_rust_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpsr_dcfr.so")
ARCH = os.environ.get("PSR_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Device code: IEEE f32/f64 semantics matching the CPU solver (no
# contraction into FMA, denormals kept), packed-FP32 instructions disabled.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]

HOST_SOURCES = ["game.cpp", "capi_game.cpp", "gpu_solver.cpp", "mlp.cpp"]
DEVICE_SOURCES = ["dcfr_kernels.hip", "mlp_kernels.hip"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        run([HIPCC, "-c", *HOST_FLAGS, f"--offload-arch={ARCH}", "-D__HIP_PLATFORM_AMD__", os.path.join(CSRC, src), "-o",
             obj])
        objs.append(obj)
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        run([HIPCC, "-c", "-x", "hip", *HOST_FLAGS, *DEVICE_FLAGS, os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
