"""Assembly-level checks of the gfx950 DCFR kernels (CPU: hipcc
cross-compiles the device code to assembly here).

  - no s_barrier inside a loop that the compiler runs under an exec mask
    (tests/isa_barrier_lint.py explains the hazard; a work-queue loop of this
    shape hung the device);
  - no scratch (private segment) in the DCFR kernels: the kernel argument block
    and per-player fields stay in registers.

The check itself is pinned on small kernels of known shape
(tests/isa_pins/barrier_loops.hip).
"""
import os
import re
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import isa_barrier_lint as L  # noqa: E402
from poker_solver_rust_amd import build as B  # noqa: E402

FLAGS = ["-std=c++17", *B.DEVICE_FLAGS]


@pytest.fixture(scope="module")
def pins_asm(tmp_path_factory):
    return L.compile_to_asm(os.path.join(HERE, "isa_pins", "barrier_loops.hip"), FLAGS,
                            str(tmp_path_factory.mktemp("pins")))


@pytest.fixture(scope="module")
def kernels_asm(tmp_path_factory):
    """Assembly of every storage mode's DCFR kernels (compiled concurrently)."""
    from concurrent.futures import ThreadPoolExecutor

    out = str(tmp_path_factory.mktemp("dcfr"))
    srcs = [os.path.join(B.CSRC, f) for f in B.DCFR_SOURCES]
    with ThreadPoolExecutor(max_workers=len(srcs)) as pool:
        return "\n".join(pool.map(lambda s: L.compile_to_asm(s, FLAGS, out), srcs))


@pytest.fixture(scope="module")
def mlp_asm(tmp_path_factory):
    return L.compile_to_asm(os.path.join(B.CSRC, "mlp_kernels.hip"), FLAGS + ["-I" + B.CSRC],
                            str(tmp_path_factory.mktemp("mlp")))


def test_mlp_kernels_have_no_barrier_in_divergent_loop(mlp_asm):
    assert not L.barriers_in_divergent_loops(mlp_asm)
    bars = L.count_barriers(mlp_asm)
    assert all(bars[f] >= 2 for f in bars if "gemm_bf16_kernel" in f)
    assert "v_mfma_f32_16x16x32_bf16" in mlp_asm


def test_check_flags_known_bad_shapes(pins_asm):
    bad = {v[0] for v in L.barriers_in_divergent_loops(pins_asm)}
    assert any("old_loop" in f for f in bad)
    assert any("lane_loop" in f for f in bad)
    assert not any("new_loop" in f for f in bad)
    assert not any("uniform_loop" in f for f in bad)
    counts = L.count_divergent_loops(pins_asm)
    assert counts[next(f for f in counts if "lane_loop" in f)] == 1


def test_dcfr_kernels_have_no_barrier_in_divergent_loop(kernels_asm):
    bad = L.barriers_in_divergent_loops(kernels_asm)
    assert not bad, bad[:5]
    # the parse found every kernel and the barriers in it
    bars = L.count_barriers(kernels_asm)
    dcfr = [f for f in bars if "dcfr_solve_kernel" in f or "dcfr_phase_kernel" in f]
    assert len(dcfr) >= 5 and all(bars[f] > 10 for f in dcfr)


def test_dcfr_kernels_use_no_scratch(kernels_asm):
    sizes = dict(re.findall(r"\.name:\s+(\S+)\s*\n(?:\s+\.[a-z_]+:.*\n)*?\s+\.private_segment_fixed_size:\s+(\d+)",
                            kernels_asm))
    dcfr = {k: int(v) for k, v in sizes.items() if "dcfr_" in k}
    assert dcfr and all(v == 0 for v in dcfr.values()), dcfr
