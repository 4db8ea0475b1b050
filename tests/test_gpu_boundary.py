"""Depth-boundary leaves on the GPU, through the C ABI, against the oracle.

Turn trees cut at the river deal (depth_limit 0): seeded leaf values, scaled
(evaluation.rs:544-604) on even ordinals and raw (evaluation.rs:606-630) on
odd ones; regrets, strategy sums, the recorded boundary reach
(evaluation.rs:91-99) and exploitability compared with the oracle after every
iteration, and the batch-wide update_leaf_cfvs (batch.rs:557-585) against
per-boundary calls.
"""
import numpy as np
import pytest

import oracle_binding as O
import poker_solver_rust_amd as P

pytestmark = pytest.mark.gpu


def boundary_game(oop="AA,KK,QQ,AKs", ip="QQ-JJ,AQs,AJs", bets=("50%, a", "2x"), turn="8d"):
    cc = P.CardConfig(range=(P.Range.parse(oop), P.Range.parse(ip)), flop=P.flop_from_str("Qs Jh 2c"),
                      turn=P.card_from_str(turn))
    s = [bets, bets]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=100, effective_stack=200, turn_bet_sizes=s,
                      river_bet_sizes=s, depth_limit=0)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def values(g, seed):
    rng = np.random.default_rng(seed)
    return [[rng.uniform(-1, 1, g.num_private_hands(q)).astype(np.float32) for q in range(2)]
            for _ in range(g.num_boundary_nodes())]


@pytest.mark.parametrize("make", [boundary_game, lambda: boundary_game(bets=("33%, 75%, a", "2.5x, a"))])
def test_boundary_stepwise_matches_oracle(make):
    g = make()
    nb = g.num_boundary_nodes()
    assert nb > 2
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 12)
    assert b.num_boundaries(0) == nb
    o = O.OracleGame(g.serialize())
    vals = values(g, 7)
    for k in range(nb):
        for q in range(2):
            b.set_boundary_cfvs(0, k, q, vals[k][q], raw=k % 2 == 1)
            o.set_boundary_cfvs(k, q, vals[k][q], raw=k % 2 == 1)
    for t in range(10):
        b.run_iterations(t, t + 1)
        o.solve_step(t)
        r, s = o.state()
        assert np.array_equal(b.regrets(0), r.astype(np.float32)), t
        assert np.array_equal(b.strategy_sum(0), s.astype(np.float32)), t
        for k in range(nb):
            for q in range(2):
                assert np.array_equal(b.boundary_reach(0, k, q), o.boundary_reach(k, q)), (t, k, q)
    assert abs(float(b.exploitability()[0]) - o.exploitability()) <= 1e-6 * max(1.0, abs(o.exploitability()))
    cf = b.root_cfvalues()[0]
    assert np.array_equal(cf[0], o.root_cfvalues(0).astype(np.float32))
    assert np.array_equal(cf[1], o.root_cfvalues(1).astype(np.float32))


def test_update_leaf_cfvs_equals_per_boundary_raw():
    """batch.rs:557-585 layout: per game, per ordinal, H_player values."""
    games = [boundary_game(), boundary_game(oop="AA,AKs", ip="KK,QQ", turn="3h")]
    a = P.GpuBatchSolver(0)
    a.prepare_batch(games, 6)
    c = P.GpuBatchSolver(0)
    c.prepare_batch(games, 6)
    p0, p1 = [], []
    for gi, g in enumerate(games):
        for k, (v0, v1) in enumerate(values(g, 11 + gi)):
            p0.append(v0)
            p1.append(v1)
            c.set_boundary_cfvs(gi, k, 0, v0, raw=True)
            c.set_boundary_cfvs(gi, k, 1, v1, raw=True)
    a.update_leaf_cfvs(np.concatenate(p0), np.concatenate(p1))
    a.run_iterations(0, 6)
    c.run_iterations(0, 6)
    for gi in range(2):
        assert np.array_equal(a.regrets(gi), c.regrets(gi))
        assert np.array_equal(a.strategy_sum(gi), c.strategy_sum(gi))
    with pytest.raises(P.PsrError):
        a.update_leaf_cfvs(np.concatenate(p0)[:-1], np.concatenate(p1))


def test_boundary_errors():
    g = boundary_game()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 4)
    with pytest.raises(P.PsrError):
        b.set_boundary_cfvs(0, g.num_boundary_nodes(), 0, np.zeros(g.num_private_hands(0), np.float32))
    with pytest.raises(P.PsrError):
        b.set_boundary_cfvs(0, 0, 0, np.zeros(g.num_private_hands(0) + 1, np.float32))
    with pytest.raises(P.PsrError):
        b.boundary_reach(1, 0, 0)
