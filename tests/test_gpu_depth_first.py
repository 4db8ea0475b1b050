"""The depth-first path of the whole-game kernel (reach and cfvalue rows in
LDS, psr_batch_set_depth_first) against the level-by-level path (rows in
HBM) on the device: the same operations in the same order, so every result
is bit-identical -- solve() stop iterations and exploitability, strategy
sums, regrets and the finalize root cfvalues.  The oracle comparisons of the
other GPU tests run river games through the depth-first path by default.
"""
import numpy as np
import pytest

import oracle_binding as O
import poker_solver_rust_amd as P
from poker_solver_rust_amd import workloads as W

pytestmark = pytest.mark.gpu


def solve_both(games, iters, targets, check=5):
    out = []
    for dfs in (True, False):
        b = P.GpuBatchSolver(0)
        b.set_depth_first(dfs)
        b.prepare_batch(games, iters, targets)
        info = b.depth_first_info()
        print(f"depth-first {dfs}: {info}", flush=True)
        e, it = b.solve(iters, check)
        out.append((b, e, it, info))
    return out


def test_river_batch_paths_identical():
    subs = W.river_batch(96, seed=20261)
    games = [s.build() for s in subs]
    (bd, ed, itd, info), (bl, el, itl, info_l) = solve_both(games, 300, [s.target for s in subs])
    assert info["games"] == len(games) and info_l["games"] == 0
    assert info["lds_bytes"] > 65536
    assert np.array_equal(itd, itl)
    assert np.array_equal(ed, el)
    for i in range(len(games)):
        assert np.array_equal(bd.strategy_sum(i), bl.strategy_sum(i)), i
        assert np.array_equal(bd.regrets(i), bl.regrets(i)), i
    rd, rl = bd.root_cfvalues(compute=False), bl.root_cfvalues(compute=False)
    for i in range(len(games)):
        for p in range(2):
            assert np.array_equal(rd[i][p], rl[i][p]), (i, p)


def test_depth_first_stepwise_matches_oracle():
    """River subgames of the benchmark distribution, state after every
    iteration against the oracle, on the depth-first path."""
    games = [s.build() for s in W.river_batch(3, seed=77, iterations=8)]
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 8)
    assert b.depth_first_info()["games"] == len(games)
    oracles = [O.OracleGame(g.serialize()) for g in games]
    for t in range(8):
        b.run_iterations(t, t + 1)
        for i, o in enumerate(oracles):
            o.solve_step(t)
            r, s = o.state()
            assert np.array_equal(b.regrets(i), r), (t, i)
            assert np.array_equal(b.strategy_sum(i), s), (t, i)
    e = b.exploitability()
    for i, o in enumerate(oracles):
        assert abs(float(e[i]) - o.exploitability()) <= 1e-6 * max(1.0, o.exploitability()), i


def test_depth_first_off_for_other_storage():
    """16-bit storage keeps the level-by-level path."""
    g = W.river_batch(1, seed=3)[0].build()
    b = P.GpuBatchSolver(0, storage="fp16")
    b.prepare_batch([g], 4)
    assert b.depth_first_info()["games"] == 0
