"""Builds tests/kernel_emu/libpsr_dcfr_emu.so: the library's own sources
(host game build, device layout, C ABI and the HIP kernels) compiled for the
CPU against hip_cpu_shim.hpp, unchanged (the shim points the kernels'
dynamic LDS block at the emulated workgroup's buffer).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "poker_solver_rust_amd", "csrc")
OUT = os.path.join(HERE, "libpsr_dcfr_emu.so")
CXX = os.environ.get("EMU_CXX", "/opt/rocm/llvm/bin/clang++")


def _inputs():
    names = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".cpp", ".hpp", ".hip", ".inc"))]
    names += [os.path.join(HERE, "hip_cpu_shim.hpp"), os.path.join(HERE, "hip", "hip_runtime.h"), __file__]
    return names


def build():
    """Builds the library unless it is newer than every source; safe to call
    from concurrent test processes (one builds under a file lock, the others
    wait and reuse its result)."""
    import fcntl

    work = os.path.join(HERE, "_build")
    os.makedirs(work, exist_ok=True)
    with open(os.path.join(work, ".lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        if os.path.exists(OUT) and os.path.getmtime(OUT) >= max(os.path.getmtime(f) for f in _inputs()):
            return OUT
        return _build(work)


def _build(work):
    flags = ["-std=c++20", "-O1", "-g", "-fPIC", "-U_FORTIFY_SOURCE", "-ffp-contract=off", "-pthread", "-I" + HERE, "-I" + CSRC,
             "-Wno-unused-function"]
    sys.path.insert(0, ROOT)
    from poker_solver_rust_amd import build as B

    srcs = [os.path.join(CSRC, f) for f in B.HOST_SOURCES]
    for f in B.DEVICE_SOURCES:   # .hip sources compiled as C++ for the CPU
        cpp = os.path.join(work, f.replace(".hip", "_emu.cpp"))
        with open(cpp, "w") as out:
            out.write(f'#include "{f}"\n')
        srcs.append(cpp)

    def compile_one(s):
        o = os.path.join(work, os.path.basename(s) + ".o")
        subprocess.check_call([CXX, "-c", *flags, s, "-o", o])
        return o

    with ThreadPoolExecutor(max_workers=max(1, min(len(srcs), os.cpu_count() or 1))) as pool:
        objs = list(pool.map(compile_one, srcs))
    tmp = OUT + ".tmp"
    subprocess.check_call([CXX, "-shared", "-pthread", *objs, "-o", tmp])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
    sys.exit(0)
