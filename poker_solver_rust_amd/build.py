"""Builds libpsr_dcfr.so (host game build + gfx950 DCFR kernels + C ABI) in
place with hipcc.  Usage: python -m poker_solver_rust_amd.build
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

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
# the DCFR kernels: one translation unit per storage mode plus the dispatch
DCFR_SOURCES = ["dcfr_kernels_f32.hip", "dcfr_kernels_ref16.hip", "dcfr_kernels_fp16.hip", "dcfr_kernels_bf16.hip"]
DEVICE_SOURCES = [*DCFR_SOURCES, "dcfr_kernels.hip", "mlp_kernels.hip", "solve_noalloc.hip"]


def run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    """Compiles every source (each its own hipcc process, run concurrently:
    the kernel file dominates), then links the shared library."""
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in DEVICE_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append((obj, [HIPCC, "-c", "-x", "hip", *HOST_FLAGS, *DEVICE_FLAGS, os.path.join(CSRC, src), "-o", obj]))
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append((obj, [HIPCC, "-c", *HOST_FLAGS, f"--offload-arch={ARCH}", "-D__HIP_PLATFORM_AMD__",
                           os.path.join(CSRC, src), "-o", obj]))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for f in [pool.submit(run, cmd) for _, cmd in jobs]:
            f.result()
    objs = [obj for obj, _ in jobs]
    tmp = LIB + ".tmp"
    run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
