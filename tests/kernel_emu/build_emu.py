"""Builds tests/kernel_emu/libpsr_dcfr_emu.so: the library's own sources
(host game build, device layout, C ABI and the HIP kernels) compiled for the
CPU against hip_cpu_shim.hpp.  The kernel source is used unchanged except
that its dynamic LDS declaration is pointed at the emulated block's buffer.
"""
import os
import re
import subprocess
import sys

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
    src = open(os.path.join(CSRC, "dcfr_kernels.hip")).read()
    src, n = re.subn(r"extern __shared__ double (\w+)\[\];", r"double* \1 = emu::dynamic_lds<double>();", src)
    assert n == 1, "dynamic LDS declaration not found"
    kern = os.path.join(work, "dcfr_kernels_emu.cpp")
    with open(kern, "w") as f:
        f.write(src)
    flags = ["-std=c++20", "-O1", "-g", "-fPIC", "-U_FORTIFY_SOURCE", "-ffp-contract=off", "-pthread", "-I" + HERE, "-I" + CSRC,
             "-Wno-unused-function"]
    objs = []
    mlp_kern = os.path.join(work, "mlp_kernels_emu.cpp")
    with open(mlp_kern, "w") as f:
        f.write(open(os.path.join(CSRC, "mlp_kernels.hip")).read())
    for s in [os.path.join(CSRC, "game.cpp"), os.path.join(CSRC, "capi_game.cpp"), os.path.join(CSRC, "gpu_solver.cpp"),
              os.path.join(CSRC, "mlp.cpp"), kern, mlp_kern]:
        o = os.path.join(work, os.path.basename(s) + ".o")
        subprocess.check_call([CXX, "-c", *flags, s, "-o", o])
        objs.append(o)
    tmp = OUT + ".tmp"
    subprocess.check_call([CXX, "-shared", "-pthread", *objs, "-o", tmp])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
    sys.exit(0)
