"""The library's DCFR kernels executed on the CPU (tests/kernel_emu: the HIP
sources compiled against a fiber-based workgroup emulation) and compared with
the oracle bit for bit.  This checks the kernels' arithmetic, indexing and
barrier structure without a GPU; the compiled gfx950 code is checked by
test_kernel_isa.py and, on the device, by the gpu-marked tests.

Each check runs in its own process (tests/kernel_emu/run_checks.py) so the
emulated library never shares a process with the HIP build; the checks are
started together when the module's first test runs (one per CPU core at a
time) and each test waits for its own.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "kernel_emu"))

import build_emu  # noqa: E402


@pytest.fixture(scope="module")
def emu_lib():
    return build_emu.build()


# (check, verify that every readfirstlane argument is wave-uniform): the
# uniformity check costs two wave exchanges per call, so the long turn
# solves run without it (the other checks cover the same code with it)
CHECKS = [("std_stepwise", True), ("std_raises_stepwise", True), ("std_solve", True), ("batch", True),
          ("compressed_stepwise", True), ("fp64_exploitability", True), ("half_storage", True),
          ("turn_stepwise", True), ("mixed_batch", False), ("boundary_stepwise", True),
          ("expected_values", False), ("raked", True), ("depth_first", True), ("wave_path", False), ("reach_at_nodes", False), ("flop_stepwise", False), ("mlp_forward", True), ("mlp_gradients", True), ("mlp_gradients_split_k", True), ("mlp_gradients_one_row", True), ("mlp_adamw_step", True),
          ("mlp_adamw_clipped", True)]


def _run_check(check, uniform):
    env = dict(os.environ)
    if not uniform:
        env["EMU_NO_UNIFORM_CHECK"] = "1"
    return subprocess.run([sys.executable, os.path.join(HERE, "kernel_emu", "run_checks.py"), check],
                          capture_output=True, text=True, timeout=900, env=env)


@pytest.fixture(scope="module")
def emu_runs(emu_lib):
    """Every check started at once on a pool of one process per core, the
    longest first; {check: future}."""
    order = ["raked", "depth_first", "wave_path", "turn_stepwise", "flop_stepwise", "mixed_batch", "expected_values"]
    todo = sorted(CHECKS, key=lambda c: order.index(c[0]) if c[0] in order else len(order))
    pool = ThreadPoolExecutor(max_workers=max(1, min(8, os.cpu_count() or 1)))
    futures = {check: pool.submit(_run_check, check, uniform) for check, uniform in todo}
    yield futures
    pool.shutdown(wait=True, cancel_futures=True)


@pytest.mark.parametrize("check,uniform", CHECKS)
def test_emulated_kernels_match_oracle(emu_runs, check, uniform):
    r = emu_runs[check].result()
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]


def test_emulation_reports_partial_barrier(tmp_path):
    """The emulation turns a barrier that only part of a workgroup reaches
    into an error instead of a hang (the failure the scheduler exists to
    catch), shown on a deliberately wrong kernel compiled for the CPU."""
    src = tmp_path / "bad.cpp"
    src.write_text('''
#include <hip/hip_runtime.h>
__global__ void bad(int* out) {
    if (threadIdx.x < 32) __syncthreads();
    out[threadIdx.x] = 1;
}
int main() {
    int* p;
    hipMalloc(&p, 256 * sizeof(int));
    hipLaunchKernelGGL(bad, dim3(1), dim3(256), 0, 0, p);
    return 0;
}
''')
    exe = tmp_path / "bad"
    subprocess.check_call([build_emu.CXX, "-std=c++20", "-O1", "-U_FORTIFY_SOURCE", "-I" + os.path.join(HERE, "kernel_emu"),
                           str(src), "-o", str(exe)])
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "cannot proceed" in r.stderr
