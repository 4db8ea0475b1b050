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
    assert info["tiles"] == len(games) and info_l["tiles"] == 0
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
    assert b.depth_first_info()["tiles"] == len(games)
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
    assert b.depth_first_info()["tiles"] == 0


def flop_game(oop, ip, board, bets="50%, a"):
    cc = P.CardConfig(range=(P.Range.parse(oop), P.Range.parse(ip)), flop=P.flop_from_str(board))
    s = [(bets, ""), (bets, "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Flop, starting_pot=100, effective_stack=100, flop_bet_sizes=s,
                      turn_bet_sizes=s, river_bet_sizes=s)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def test_subtree_tiles_paths_identical():
    """Turn and flop games: their river subtree tiles run depth first in the
    subtree phase (root reach from the chance node's HBM row, root cfvalues
    back to HBM for the chance sums), the flop games' one tile per wave and
    the turn game's one per workgroup, in the same launch sequence; solve(), download_reach after the
    iterations and after a best response, and reach_at_nodes equal to the
    level-by-level path."""
    games = [W.turn_uniform01(20260).build(), flop_game("AA,KK,AKs", "QQ-JJ,AQs", "Qs Jh 2c"),
             flop_game("AA,KK", "QQ,AK", "2s 7s Js", "a")]
    runs = []
    for dfs in (True, False):
        b = P.GpuBatchSolver(0)
        b.set_depth_first(dfs)
        b.prepare_batch(games, 30, [0.001] * len(games))
        info = b.depth_first_info()
        print(f"depth-first {dfs}: {info}", flush=True)
        assert (info["tiles"] > 0) == dfs
        # the flop games' subtree tiles run one per wave, the U[0,1] turn game's per workgroup
        assert (0 < info["wave_tiles"] < info["tiles"]) if dfs else info["wave_tiles"] == 0
        b.run_iterations(0, 7)
        r_iter = b.download_reach(1326)
        e = b.exploitability()
        r_br = b.download_reach(1326)
        reach = b.reach_at_nodes(1, 0, list(range(0, games[1].num_nodes(), 97)))
        b.reset()
        es, its = b.solve(30, 5)
        runs.append((b, e, r_iter, r_br, reach, es, its))
    (bd, *d), (bl, *l) = runs
    for x, y, what in zip(d, l, ["exploitability", "download_reach", "download_reach after br", "reach_at_nodes",
                                 "solve exploitability", "solve iterations"]):
        assert np.array_equal(x, y), what
    for i in range(len(games)):
        assert np.array_equal(bd.strategy_sum(i), bl.strategy_sum(i)), i
        assert np.array_equal(bd.regrets(i), bl.regrets(i)), i
        for p in range(2):
            assert np.array_equal(bd.root_cfvalues(compute=False)[i][p], bl.root_cfvalues(compute=False)[i][p]), (i, p)


def test_wave_path_raked_flop():
    """A raked flop game (rake 5%, cap 30): its river tiles run one per wave
    through the raked showdown (win, tie and lose bands); state after every
    iteration equal to the oracle's and to the level-by-level path, then the
    exploitability (the best response and the average strategy's value)."""
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs"), P.Range.parse("QQ-TT,AQs,KQs")),
                      flop=P.flop_from_str("Qs Jh 2c"))
    s = [("50%, a", ""), ("50%, a", "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Flop, starting_pot=100, effective_stack=100, flop_bet_sizes=s,
                      turn_bet_sizes=s, river_bet_sizes=s, rake_rate=0.05, rake_cap=30.0)
    g = P.PostFlopGame.with_config(cc, P.ActionTree(tc))
    runs = []
    for dfs in (True, False):
        b = P.GpuBatchSolver(0)
        b.set_depth_first(dfs)
        b.prepare_batch([g], 4)
        info = b.depth_first_info()
        assert (info["wave_tiles"] > 0) == dfs, info
        runs.append(b)
    b, lv = runs
    o = O.OracleGame(g.serialize())
    for t in range(3):
        b.run_iterations(t, t + 1)
        lv.run_iterations(t, t + 1)
        o.solve_step(t)
        r, st = o.state()
        assert np.array_equal(b.regrets(0), r), t
        assert np.array_equal(b.strategy_sum(0), st), t
        assert np.array_equal(b.regrets(0), lv.regrets(0)), t
    e, el = b.exploitability(), lv.exploitability()
    assert np.array_equal(e, el)
    assert abs(float(e[0]) - o.exploitability()) <= 1e-6 * max(1.0, o.exploitability())
