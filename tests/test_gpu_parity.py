"""Parity of the HIP DCFR path (through the C ABI) with the CPU oracle.

Bar: per-hand CFVs, regrets, strategy sums and exploitability within 1e-4
relative to the vector's largest magnitude (f32 values; the device uses the
reference operation order, so in practice they agree bit for bit), the
reference solve() schedule (iterations run, early stop) identical.  Full-size
batches are checked through size-independent properties.
"""
import numpy as np
import pytest

import oracle_binding as O
import poker_solver_rust_amd as P
from poker_solver_rust_amd import workloads as W

pytestmark = pytest.mark.gpu
RTOL = 1e-4


def close(gpu, ref, rtol=RTOL):
    gpu = np.asarray(gpu, np.float64)
    ref = np.asarray(ref, np.float64)
    assert gpu.shape == ref.shape
    scale = max(np.abs(ref).max(), 1e-30)
    err = np.abs(gpu - ref).max() / scale
    assert err <= rtol, f"max rel err {err:.3e} ({int((gpu != ref).sum())} of {gpu.size} differ)"


def std_game(bets="50%, a", raises=""):
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs"), P.Range.parse("QQ-JJ,AQs,AJs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"), river=P.card_from_str("3s"))
    tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=100,
                      river_bet_sizes=[(bets, raises), (bets, raises)])
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def test_stepwise_state_matches_oracle():
    g = std_game()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 20)
    o = O.OracleGame(g.serialize())
    for t in range(12):
        b.run_iterations(t, t + 1)
        o.solve_step(t)
        r, s = o.state()
        close(b.regrets(0), r)
        close(b.strategy_sum(0), s)
    assert abs(float(b.exploitability()[0]) - o.exploitability()) <= RTOL * max(1.0, o.exploitability())
    cf = b.root_cfvalues()[0]
    close(cf[0], o.root_cfvalues(0))
    close(cf[1], o.root_cfvalues(1))


def test_solve_schedule_matches_oracle():
    g = std_game("33%, 75%, a", "2.5x")
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 200, [0.5])
    e, it = b.solve(200, 5)
    o = O.OracleGame(g.serialize())
    eo, ito = o.solve(200, 0.5)
    assert int(it[0]) == ito
    assert abs(float(e[0]) - eo) <= RTOL * max(1.0, eo)
    r, s = o.state()
    close(b.regrets(0), r)
    close(b.strategy_sum(0), s)


def test_gpu_solve_hand_parallel_api():
    """gpu-range-solver lib.rs:214-245: exploitability >= ~0, < 1.0 after 500.
    The stop test is `exploitability <= target` (gpu-range-solver
    solver.rs:635, range-solver solver.rs:137-139), so with target 0 a game
    whose exploitability reaches exactly 0.0 stops early, on the oracle too."""
    g = std_game("100%")
    res = P.gpu_solve_hand_parallel(g, P.GpuSolverConfig(max_iterations=500, target_exploitability=0.0))
    o = O.OracleGame(g.serialize())
    eo, ito = o.solve(500)
    assert -1e-4 < res.exploitability < 1.0 and res.iterations_run == ito
    assert abs(res.exploitability - eo) < 0.01
    A, H = 2, g.num_private_hands(0)
    strat = res.root_strategy[:A * H].reshape(A, H)
    assert np.allclose(strat.sum(0), 1.0, atol=1e-6)


def test_river_uniform_1000_iterations():
    """1000-iteration river subgame, 1081 hands per player."""
    s = W.river_uniform(0)
    g = s.build()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 1000)
    e, it = b.solve(1000, 5)
    o = O.OracleGame(g.serialize())
    eo, ito = o.solve(1000)
    assert int(it[0]) == ito == 1000
    assert abs(float(e[0]) - eo) <= RTOL * max(1.0, eo)
    r, st = o.state()
    close(b.strategy_sum(0), st)
    close(b.regrets(0), r)
    cf = b.root_cfvalues()[0]
    close(cf[0], o.root_cfvalues(0))
    close(cf[1], o.root_cfvalues(1))
    assert float(e[0]) <= 0.003 * s.pot


def test_river_batch_matches_oracle():
    """Heterogeneous batch: every game agrees with its own oracle solve."""
    subs = W.river_batch(12, seed=11, iterations=300)
    games = [x.build() for x in subs]
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 300, [x.target for x in subs])
    e, it = b.solve(300, 5)
    cfs = b.root_cfvalues()
    for i, g in enumerate(games):
        o = O.OracleGame(g.serialize())
        eo, ito = o.solve(300, subs[i].target)
        assert int(it[i]) == ito, i
        assert abs(float(e[i]) - eo) <= RTOL * max(1.0, abs(eo)), i
        r, st = o.state()
        close(b.strategy_sum(i), st)
        close(cfs[i][0], o.root_cfvalues(0))
        close(cfs[i][1], o.root_cfvalues(1))


def test_batch_equals_single():
    subs = W.river_batch(6, seed=12, iterations=50)
    games = [x.build() for x in subs]
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 50)
    b.run_iterations(0, 50)
    single = P.GpuBatchSolver(0)
    single.prepare_batch([games[3]], 50)
    single.run_iterations(0, 50)
    assert np.array_equal(b.strategy_sum(3), single.strategy_sum(0))
    assert np.array_equal(b.regrets(3), single.regrets(0))


def test_incremental_iterations_equal_one_launch():
    """batch.rs:411-534: state persists across run_iterations calls."""
    g = std_game()
    a = P.GpuBatchSolver(0)
    a.prepare_batch([g], 40)
    a.run_iterations(0, 40)
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 40)
    for s in range(0, 40, 7):
        b.run_iterations(s, min(s + 7, 40))
    assert np.array_equal(a.regrets(0), b.regrets(0))
    b.reset()
    assert not b.regrets(0).any() and not b.strategy_sum(0).any()


def test_edge_games():
    """Tiny stacks (all-in only), single-hand ranges and zero-weight combos."""
    a = np.zeros(1326, np.float32)
    c = np.zeros(1326, np.float32)
    a[P.card_pair_to_index(P.card_from_str("Ah"), P.card_from_str("Ad"))] = 1.0
    c[P.card_pair_to_index(P.card_from_str("Kh"), P.card_from_str("Kd"))] = 1.0
    c[P.card_pair_to_index(P.card_from_str("Qh"), P.card_from_str("Qd"))] = 0.25
    games = []
    for stack, bets in [(1, "50%"), (100, "50%, a"), (1000, "25%, 75%, 150%, a")]:
        cc = P.CardConfig(range=(P.Range(a), P.Range(c)), flop=P.flop_from_str("2c7s9h"),
                          turn=P.card_from_str("Js"), river=P.card_from_str("3c"))
        tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=stack,
                          river_bet_sizes=[(bets, "2x, a"), (bets, "2x, a")])
        games.append(P.PostFlopGame.with_config(cc, P.ActionTree(tc)))
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 100)
    e, it = b.solve(100, 5)
    for i, g in enumerate(games):
        o = O.OracleGame(g.serialize())
        eo, ito = o.solve(100)
        assert int(it[i]) == ito and abs(float(e[i]) - eo) <= 1e-5
        close(b.strategy_sum(i), o.state()[1])


def test_full_size_batch_properties():
    """4096 seeded subgames at full size: exploitability >= 0, every game at
    or near 0.3% pot, and the unraked zero-sum identity of the root values
    under the average strategy (sum_h w_h * cfv_h over both players = 0)."""
    subs = W.river_batch(4096, seed=20260, iterations=300)
    games = [x.build() for x in subs]
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 300, [x.target for x in subs])
    e, it = b.solve(300, 5)
    pots = np.array([x.pot for x in subs], np.float64)
    assert np.all(e >= -1e-4 * pots)
    assert np.mean(e <= 0.003 * pots + 1e-6) >= 0.99
    cfs = b.root_cfvalues()
    for i in range(0, 4096, 97):
        w0, w1 = games[i].initial_weights(0), games[i].initial_weights(1)
        v0 = float(np.dot(cfs[i][0].astype(np.float64), w0))
        v1 = float(np.dot(cfs[i][1].astype(np.float64), w1))
        assert abs(v0 + v1) <= 1e-4 * max(abs(v0), abs(v1), 1e-3 * pots[i]), i


def test_error_paths():
    b = P.GpuBatchSolver(0)
    with pytest.raises(P.PsrError):
        b.run_iterations(0, 1)           # before prepare_batch (batch.rs:419-422)
    b.prepare_batch([std_game()], 10)
    with pytest.raises(P.PsrError):
        b.run_iterations(0, 11)          # beyond the prepared iterations
    with pytest.raises(P.PsrError):
        b.solve(11, 5)


# ------------------------------------------------------------------ turn --

def turn_game(oop="AA,KK,QQ,AKs", ip="QQ-JJ,AQs,AJs", board="Qs Jh 2c", turn="8d", bets="50%, a", raises="",
              pot=100, stack=200, max_raises=0):
    cc = P.CardConfig(range=(P.Range.parse(oop), P.Range.parse(ip)), flop=P.flop_from_str(board),
                      turn=P.card_from_str(turn))
    s = [(bets, raises), (bets, raises)]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=pot, effective_stack=stack, turn_bet_sizes=s,
                      river_bet_sizes=s, max_raises=max_raises)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def test_turn_stepwise_state_matches_oracle():
    g = turn_game()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 20)
    o = O.OracleGame(g.serialize())
    for t in range(8):
        b.run_iterations(t, t + 1)
        o.solve_step(t)
        r, s = o.state()
        close(b.regrets(0), r)
        close(b.strategy_sum(0), s)
    assert abs(float(b.exploitability()[0]) - o.exploitability()) <= RTOL * max(1.0, o.exploitability())
    cf = b.root_cfvalues()[0]
    close(cf[0], o.root_cfvalues(0))
    close(cf[1], o.root_cfvalues(1))


def test_turn_suit_isomorphic_rivers():
    """No clubs or diamonds on the board and suit-symmetric ranges: diamond
    rivers are skipped and added back through the swap lists
    (solver.rs:307-320, isomorphism.rs)."""
    g = turn_game(oop="AA,KK,QQ,AKs,T9s", ip="QQ-TT,AQs,AJs,87s", board="Ah Kh Qs", turn="Js")
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 40, [0.0])
    e, it = b.solve(40, 5)
    o = O.OracleGame(g.serialize())
    eo, ito = o.solve(40)
    assert int(it[0]) == ito
    assert abs(float(e[0]) - eo) <= RTOL * max(1.0, eo)
    close(b.strategy_sum(0), o.state()[1])
    cf = b.root_cfvalues()[0]
    close(cf[0], o.root_cfvalues(0))


def test_turn_basic_example_game():
    """examples/basic.rs turn game (250 IP hands), solve() schedule."""
    oop = P.Range.parse("66+,A8s+,A5s-A4s,AJo+,K9s+,KQo,QTs+,JTs,96s+,85s+,75s+,65s,54s")
    ip = P.Range.parse("QQ-22,AQs-A2s,ATo+,K5s+,KJo+,Q8s+,J8s+,T7s+,96s+,86s+,75s+,64s+,53s+")
    cc = P.CardConfig(range=(oop, ip), flop=P.flop_from_str("Td9d6h"), turn=P.card_from_str("Qc"))
    s = [("60%, e, a", "2.5x"), ("60%, e, a", "2.5x")]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=200, effective_stack=900, flop_bet_sizes=s,
                      turn_bet_sizes=s, river_bet_sizes=s, add_allin_threshold=1.5, force_allin_threshold=0.15,
                      merging_threshold=0.1)
    g = P.PostFlopGame.with_config(cc, P.ActionTree(tc))
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 30, [1.0])
    e, it = b.solve(30, 5)
    o = O.OracleGame(g.serialize())
    eo, ito = o.solve(30, 1.0)
    assert int(it[0]) == ito
    assert abs(float(e[0]) - eo) <= RTOL * max(1.0, eo)
    close(b.strategy_sum(0), o.state()[1])


def test_turn_full_size_iterations_match_oracle():
    """The 500-iteration turn config at full size (1128 hands, ~1.7e4 nodes):
    the first iterations agree with the oracle state element for element."""
    s = W.turn_uniform01(0)
    g = s.build()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 500)
    b.run_iterations(0, 3)
    o = O.OracleGame(g.serialize())
    for t in range(3):
        o.solve_step(t)
    r, st = o.state()
    close(b.regrets(0), r)
    close(b.strategy_sum(0), st)


def test_mixed_river_and_turn_batch():
    """River (whole-game kernel) and turn (phase launches) games in one batch."""
    games = [std_game(), turn_game(), W.river_batch(1, seed=5, iterations=30)[0].build()]
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 30)
    e, it = b.solve(30, 5)
    for i, g in enumerate(games):
        o = O.OracleGame(g.serialize())
        eo, ito = o.solve(30)
        assert int(it[i]) == ito
        assert abs(float(e[i]) - eo) <= RTOL * max(1.0, eo)
        close(b.strategy_sum(i), o.state()[1])


def test_mixed_batch_equals_separate_batches():
    """The whole-game kernel of a mixed batch runs on a second stream beside
    the turn game's phase loop: every result equals that of the same games
    solved in batches of one kind (solve() iterations and exploitability,
    strategy sums, regrets)."""
    rivers = [s.build() for s in W.river_batch(24, seed=31, iterations=30)]
    turn = turn_game()
    targets = [0.005] * 25

    def run(games, tg):
        b = P.GpuBatchSolver(0)
        b.prepare_batch(games, 30, tg)
        e, it = b.solve(30, 5)
        return b, e, it

    bm, em, im = run(rivers + [turn], targets)
    br, er, ir = run(rivers, targets[:24])
    bt, et, it_ = run([turn], targets[:1])
    assert np.array_equal(em[:24], er) and np.array_equal(im[:24], ir)
    assert np.array_equal(em[24:], et) and np.array_equal(im[24:], it_)
    for i in range(24):
        assert np.array_equal(bm.strategy_sum(i), br.strategy_sum(i)), i
        assert np.array_equal(bm.regrets(i), br.regrets(i)), i
    assert np.array_equal(bm.strategy_sum(24), bt.strategy_sum(0))
    assert np.array_equal(bm.regrets(24), bt.regrets(0))


@pytest.mark.parametrize("make,iters", [(lambda: std_game("33%, 75%, a", "2x, a"), 40),
                                        (lambda: W.river_uniform(2).build(), 60), (turn_game, 10),
                                        (lambda: flop_game("AA,KK", "QQ,AKs", "Qs Jh 2c", "50%, a"), 3)])
def test_expected_values_match_oracle(make, iters):
    """query.rs:498-760 at the root: normalized weights, per-action detail of
    the acting player and expected values of both players."""
    g = make()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], iters)
    b.run_iterations(0, iters)
    o = O.OracleGame(g.serialize())
    for t in range(iters):
        o.solve_step(t)
    for p in range(2):
        assert np.array_equal(b.normalized_weights(0, p), o.normalized_weights(p))
        close(b.expected_values_detail(0, p), o.expected_values(p, detail=True))
        close(b.expected_values(0, p), o.expected_values(p))


# ----------------------------------------------------------------- raked --

def raked_river(bets="50%, a", raises="", rake=(0.05, 3.0), seed=None):
    if seed is not None:
        s = W.river_batch(1, seed=seed, iterations=30)[0]
        tc = s.tree_config()
        tc.rake_rate, tc.rake_cap = rake
        return P.PostFlopGame.with_config(s.card_config(), P.ActionTree(tc))
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs"), P.Range.parse("QQ-JJ,AQs,AJs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"), river=P.card_from_str("3s"))
    tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=100,
                      river_bet_sizes=[(bets, raises), (bets, raises)], rake_rate=rake[0], rake_cap=rake[1])
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def raked_turn(rake=(0.05, 3.0)):
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs"), P.Range.parse("QQ-JJ,AQs,AJs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"))
    s = [("50%, a", ""), ("50%, a", "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=100, effective_stack=200, turn_bet_sizes=s,
                      river_bet_sizes=s, rake_rate=rake[0], rake_cap=rake[1])
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


@pytest.mark.parametrize("make,iters", [(raked_river, 10), (lambda: raked_river("33%, 75%, a", "2x, a", (0.1, 1e9)), 6),
                                        (raked_turn, 4)])
def test_raked_stepwise_matches_oracle(make, iters):
    """Raked showdown values (evaluation.rs:429-537) and the raked
    exploitability (utility.rs:380-392: best response minus the value of the
    average strategy), iteration by iteration; then the f64 best response."""
    g = make()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], iters)
    o = O.OracleGame(g.serialize())
    for t in range(iters):
        b.run_iterations(t, t + 1)
        o.solve_step(t)
        r, s = o.state()
        close(b.regrets(0), r)
        close(b.strategy_sum(0), s)
    assert abs(float(b.exploitability()[0]) - o.exploitability()) <= RTOL * max(1.0, o.exploitability())
    cf = b.root_cfvalues()[0]
    close(cf[0], o.root_cfvalues(0))
    close(cf[1], o.root_cfvalues(1))
    o64 = O.OracleGame(g.serialize(), 64)
    o64.set_state(b.regrets(0).astype(np.float64), b.strategy_sum(0).astype(np.float64))
    e64 = float(b.exploitability_f64()[0])
    assert abs(e64 - o64.exploitability()) <= 1e-9 * max(1.0, abs(o64.exploitability()))


def test_raked_and_unraked_batch_solve():
    """solve() schedule on a batch mixing raked and unraked river and turn
    games: each game's iterations and exploitability as its own oracle's."""
    games = [raked_river(), std_game(), raked_river(seed=11), raked_turn(), turn_game()]
    b = P.GpuBatchSolver(0)
    targets = [0.5, 0.5, 0.0, 0.5, 0.5]
    b.prepare_batch(games, 40, targets)
    e, it = b.solve(40, 5)
    for i, g in enumerate(games):
        eo, ito = O.OracleGame(g.serialize()).solve(40, targets[i])
        assert int(it[i]) == ito, i
        assert abs(float(e[i]) - eo) <= RTOL * max(1.0, eo), i


# ------------------------------------------------------ reach at nodes --

@pytest.mark.parametrize("make,iters", [(lambda: std_game("33%, 75%, a", "2x, a"), 20), (turn_game, 6),
                                        (lambda: W.river_batch(1, seed=9, iterations=30)[0].build(), 15),
                                        (lambda: flop_game("AA,KK", "QQ,AKs", "Qs Jh 2c", "50%, a"), 3)])
def test_reach_at_nodes_matches_oracle(make, iters):
    """batch.rs:765-797 compute_reach_at_nodes: the reach of each player at
    every node under the average strategy, against the oracle's walk."""
    g = make()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], iters)
    b.run_iterations(0, iters)
    o = O.OracleGame(g.serialize())
    o.set_state(b.regrets(0).astype(np.float64), b.strategy_sum(0).astype(np.float64))
    nodes = np.arange(g.num_nodes())
    rng = np.random.default_rng(1)
    sample = nodes if len(nodes) <= 400 else np.sort(rng.choice(nodes, 400, replace=False))
    for p in range(2):
        got = b.reach_at_nodes(0, p, sample)
        for k, n in enumerate(sample):
            close(got[k], o.reach_at_node(int(n), p))
    with pytest.raises(P.PsrError):
        b.reach_at_nodes(0, 0, [g.num_nodes()])


# ------------------------------------------------------------------ flop --

def flop_game(oop, ip, board, bets="a", raises=""):
    cc = P.CardConfig(range=(P.Range.parse(oop), P.Range.parse(ip)), flop=P.flop_from_str(board))
    s = [(bets, raises), (bets, raises)]
    tc = P.TreeConfig(initial_state=P.BoardState.Flop, starting_pot=100, effective_stack=100, flop_bet_sizes=s,
                      turn_bet_sizes=s, river_bet_sizes=s)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


@pytest.mark.parametrize("make", [lambda: flop_game("AA,KK", "QQ,AKs", "Qs Jh 2c"),
                                  lambda: flop_game("AA,KK", "QQ,AK", "2s 7s Js"),
                                  lambda: flop_game("AA,KK,AKs", "QQ-JJ,AQs", "Qs Jh 2c", "50%, a")])
def test_flop_stepwise_matches_oracle(make):
    """Flop-root games: turn subtrees as tiles holding their river deals
    (chance nodes inside a tile; suit-isomorphic turn and river cards on the
    monotone flop), state after every iteration, exploitability in f32 and
    f64, root cfvalues."""
    g = make()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 4)
    o = O.OracleGame(g.serialize())
    for t in range(3):
        b.run_iterations(t, t + 1)
        o.solve_step(t)
        r, s = o.state()
        close(b.regrets(0), r)
        close(b.strategy_sum(0), s)
    assert abs(float(b.exploitability()[0]) - o.exploitability()) <= RTOL * max(1.0, o.exploitability())
    cf = b.root_cfvalues()[0]
    close(cf[0], o.root_cfvalues(0))
    close(cf[1], o.root_cfvalues(1))
    o64 = O.OracleGame(g.serialize(), 64)
    o64.set_state(b.regrets(0).astype(np.float64), b.strategy_sum(0).astype(np.float64))
    assert abs(float(b.exploitability_f64()[0]) - o64.exploitability()) <= 1e-9 * max(1.0, abs(o64.exploitability()))


def test_flop_solve_schedule_matches_oracle():
    g = flop_game("AA,KK", "QQ,AK", "2s 7s Js")
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 40, [0.5])
    e, it = b.solve(40, 5)
    eo, ito = O.OracleGame(g.serialize()).solve(40, 0.5)
    assert int(it[0]) == ito
    assert abs(float(e[0]) - eo) <= RTOL * max(1.0, eo)


def test_several_multi_tile_games_in_one_batch():
    """Two turn games, a flop game and a river game in one batch: the root
    tiles' chance nodes and terminals of every multi-tile game share one
    parallel phase.  State after each iteration, then the solve() schedule
    with targets that stop the games at different iterations (a stopped
    game's phase work is skipped)."""
    games = [turn_game(), turn_game("KK,QQ,AKo", "JJ-99,KQs", "Ks 7d 2h", "4s", "33%, a", "2x"),
             flop_game("AA,KK", "QQ,AKs", "Qs Jh 2c"), std_game()]
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 4)
    oracles = [O.OracleGame(g.serialize()) for g in games]
    for t in range(3):
        b.run_iterations(t, t + 1)
        for i, o in enumerate(oracles):
            o.solve_step(t)
            r, s = o.state()
            close(b.regrets(i), r)
            close(b.strategy_sum(i), s)
    targets = [2.0, 0.5, 1.0, 0.3]
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 40, targets)
    e, it = b.solve(40, 5)
    for i, g in enumerate(games):
        eo, ito = O.OracleGame(g.serialize()).solve(40, targets[i])
        assert int(it[i]) == ito
        assert abs(float(e[i]) - eo) <= RTOL * max(1.0, eo)
