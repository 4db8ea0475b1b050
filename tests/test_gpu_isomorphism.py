"""Suit isomorphism on the device: a turn game and a flop game whose ranges
and boards are symmetric under a suit swap, solved on the GPU with the
isomorphic chance sums (cards skipped, hand swaps, isomorphism.rs:33-237)
and without them (every card dealt).  The f32 DCFR trajectories are
identical (the f64 chance sums of the dealt cards are exact in any order),
so the root CFVs, exploitability and strategy sums at the root must agree
exactly; the device's isomorphic game is also compared with the oracle."""
import numpy as np
import pytest

import oracle_binding as O
import poker_solver_rust_amd as P
from test_isomorphism import FLOP, TURN, build

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case,iters", [(TURN, 40), (FLOP, 12)], ids=["turn", "flop"])
def test_device_isomorphic_chance_sums_equal_expanded_game(case, iters):
    gi, gn = build(case, True), build(case, False)
    out = []
    for g in (gi, gn):
        b = P.GpuBatchSolver(0)
        b.prepare_batch([g], iters)
        b.run_iterations(0, iters)
        cf = b.root_cfvalues()[0]
        a_root = len(b.expected_values_detail(0, 0)) // g.num_private_hands(0)
        root_s = b.strategy_sum(0)[:a_root * g.num_private_hands(0)]
        out.append((float(b.exploitability()[0]), cf, root_s))
    (ei, ci, si), (en, cn, sn) = out
    assert ei == en
    for p in range(2):
        assert np.array_equal(ci[p], cn[p])
    assert np.array_equal(si, sn)
    o = O.OracleGame(gi.serialize())
    for t in range(iters):
        o.solve_step(t)
    assert abs(ei - o.exploitability()) <= 1e-4 * max(1.0, o.exploitability())
