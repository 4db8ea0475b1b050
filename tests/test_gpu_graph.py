"""Replayed hipGraphs of the multi-tile phase launches and the launch-shape
autotuner, on the device.  The graph of a block of DCFR iterations reads its
iteration index from device memory and advances it itself; a block never
spans an exploitability check.  Every launch shape gives the same regrets,
strategy sums, iteration counts and exploitability as launches issued one by
one from the host."""
import numpy as np
import pytest

import poker_solver_rust_amd as P
from poker_solver_rust_amd import workloads as W

pytestmark = pytest.mark.gpu


def turn_game():
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs,AQo"), P.Range.parse("QQ-99,AJs,KQs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"))
    s = [("50%, a", "2x"), ("50%, a", "2x")]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=100, effective_stack=200, turn_bet_sizes=s,
                      river_bet_sizes=s)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def flop_game():
    cc = P.CardConfig(range=(P.Range.parse("AA,KK"), P.Range.parse("QQ,AKs")), flop=P.flop_from_str("Qs Jh 2c"))
    s = [("50%, a", ""), ("50%, a", "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Flop, starting_pot=100, effective_stack=100, flop_bet_sizes=s,
                      turn_bet_sizes=s, river_bet_sizes=s)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def run(games, shape, iters, solve):
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, iters, [0.05] * len(games) if solve else None)
    b.set_launch(*shape)
    if solve:
        e, it = b.solve(iters, 5)
    else:
        b.run_iterations(0, 7)
        b.run_iterations(7, iters)
        e, it = b.exploitability(), None
    return b, e, it


SHAPES = [(1, 1, 1), (5, 1, 1), (5, 2, 1), (3, 2, 1)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "graph%d_subtree%d" % s[:2])
@pytest.mark.parametrize("solve", [False, True], ids=["run_iterations", "solve"])
def test_graph_replay_equals_eager_launches(shape, solve):
    games = [turn_game(), flop_game(), W.river_uniform(3).build()]
    ref, e0, it0 = run(games, (0, 1, 1), 23, solve)
    got, e1, it1 = run(games, shape, 23, solve)
    assert np.array_equal(e0, e1)
    if solve:
        assert np.array_equal(it0, it1)
    for i in range(len(games)):
        assert np.array_equal(ref.regrets(i), got.regrets(i))
        assert np.array_equal(ref.strategy_sum(i), got.strategy_sum(i))


def test_autotune_picks_a_shape_and_caches_it():
    games = [turn_game(), turn_game()]
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 40)
    shape = b.autotune(6)
    assert not shape["cached"] and shape["graph_iterations"] in (0, 1, 5) and shape["subtree_per_cu"] in (1, 2)
    # the batch is left reset: the tuned solve equals a fresh eager one
    b.run_iterations(0, 12)
    ref = P.GpuBatchSolver(0)
    ref.prepare_batch(games, 40)
    ref.set_launch(0, 1, 1)
    ref.run_iterations(0, 12)
    assert np.array_equal(b.strategy_sum(0), ref.strategy_sum(0))
    again = P.GpuBatchSolver(0)
    again.prepare_batch(games, 40)
    assert again.autotune(6)["cached"]


def test_set_launch_rejects_bad_shapes():
    b = P.GpuBatchSolver(0)
    b.prepare_batch([turn_game()], 5)
    for bad in [(-1, 1, 1), (5, 0, 1), (5, 1, 0), (5, 1, 99)]:
        with pytest.raises(P.PsrError):
            b.set_launch(*bad)
