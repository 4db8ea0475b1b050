"""The reference's own known-answer tests for regret matching and the slice
operations, run on the oracle's functions (the ones its DCFR recursion is
written with):
crates/range-solver/src/solver.rs:724-770 (regret matching, plain and
compressed), crates/range-solver/src/sliceop.rs:195-272 (slice ops)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import oracle_binding as O  # noqa: E402


def test_regret_matching_all_negative():
    s = O.regret_matching([-1.0, -2.0, -3.0, -1.0, -2.0, -3.0], 3)
    assert np.all(np.abs(s - 1.0 / 3.0) < 1e-6)


def test_regret_matching_positive():
    s = O.regret_matching([1.0, 0.0, 0.0, 3.0], 2)
    assert np.allclose(s, [1.0, 0.0, 0.0, 1.0], atol=1e-6)


def test_regret_matching_mixed():
    s = O.regret_matching([2.0, 1.0], 2)
    assert abs(s[0] - 2.0 / 3.0) < 1e-6 and abs(s[1] - 1.0 / 3.0) < 1e-6


def test_regret_matching_compressed_all_negative():
    s = O.regret_matching_compressed([-100, -200, -300, -100, -200, -300], 3)
    assert np.all(np.abs(s - 1.0 / 3.0) < 1e-6)


def test_regret_matching_compressed_positive():
    s = O.regret_matching_compressed([100, 0, 0, 200], 2)
    assert np.allclose(s, [1.0, 0.0, 0.0, 1.0], atol=1e-6)


def test_regret_matching_rows_are_distributions():
    """Beyond the reference's cases: every hand's column sums to 1 within
    f32 rounding, and equals max(R, 0) / sum wherever the sum is positive."""
    rng = np.random.default_rng(5)
    for A in (1, 2, 3, 5):
        r = rng.normal(0, 1, A * 37).astype(np.float32)
        r[:37] = -np.abs(r[:37])   # some hands have every regret negative
        s = O.regret_matching(r, A).reshape(A, 37)
        assert np.allclose(s.sum(0), 1.0, atol=1e-6)
        pos = np.maximum(r.reshape(A, 37), 0)
        den = pos.sum(0)
        live = den > 0
        assert np.allclose(s[:, live], pos[:, live] / den[live], rtol=1e-6)
        assert np.allclose(s[:, ~live], 1.0 / A)


def test_max():
    assert O.slice_op("max", 1.0, 2.0) == 2.0
    assert O.slice_op("max", -1.0, -2.0) == -1.0
    assert O.slice_op("max", 0.0, 0.0) == 0.0


def test_sub_slice():
    assert O.slice_op("sub", [3.0, 5.0, 7.0], [1.0, 2.0, 3.0]).tolist() == [2.0, 3.0, 4.0]


def test_mul_slice():
    assert O.slice_op("mul", [2.0, 3.0, 4.0], [1.0, 2.0, 0.5]).tolist() == [2.0, 6.0, 2.0]


def test_div_slice():
    assert O.slice_op("div", [6.0, 3.0, 0.0], [2.0, 0.0, 5.0], default=99.0).tolist() == [3.0, 99.0, 0.0]


def test_div_slice_negative_zero_is_zero():
    """div_slice tests `r == 0.0`, which holds for -0.0 too (sliceop.rs:35-39)."""
    assert O.slice_op("div", [1.0], [-0.0], default=7.0).tolist() == [7.0]


def test_sum_slices_uninit():
    # 3 actions, 2 hands => src = [a0h0, a0h1, a1h0, a1h1, a2h0, a2h1]
    assert O.slice_op("sum", None, [1.0, 2.0, 3.0, 4.0, 5.0, 6.0], row=2).tolist() == [9.0, 12.0]


def test_fma_slices_uninit():
    r = O.slice_op("fma", None, [0.5, 0.5, 0.5, 0.5], [10.0, 20.0, 30.0, 40.0], row=2)
    assert r.tolist() == [20.0, 30.0]


def test_max_slices_uninit():
    r = O.slice_op("max_slices", None, [1.0, 5.0, 4.0, 2.0, -1.0, 9.0], row=2)
    assert r.tolist() == [4.0, 9.0]


@pytest.mark.parametrize("op", ["sum", "fma"])
def test_row_reductions_sum_in_row_order(op):
    """f32 sums in row order: (a0 + a1) + a2, which differs from other
    association orders on these values."""
    rows = np.array([1e8, 0.0, 1.0, 0.0, -1e8, 0.0], np.float32)   # 3 rows of 2 values
    one = np.ones_like(rows)
    r = O.slice_op(op, None, rows, one if op == "fma" else None, row=2)
    # sequential f32: (1e8 + 1) = 1e8, then - 1e8 = 0 (1e8 + (1 - 1e8) would give 8)
    assert r[0] == np.float32(0.0)
