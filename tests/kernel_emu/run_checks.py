"""Runs the library's DCFR kernels on the CPU (libpsr_dcfr_emu.so, built by
build_emu.py) against the oracle.  Started in its own process by
tests/test_kernel_emu.py so that the emulated library never shares a process
with the HIP one.  Usage: run_checks.py CHECK [CHECK ...]; exit status 0 when
every check passes, the failing check named on stderr otherwise.
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

import oracle_binding as O  # noqa: E402
import poker_solver_rust_amd as P  # noqa: E402
from poker_solver_rust_amd import range_solver as R  # noqa: E402
from poker_solver_rust_amd import workloads as W  # noqa: E402

R._LIB = R.load_library(os.path.join(HERE, "libpsr_dcfr_emu.so"))


def std_game(bets="50%, a", raises="", rake=(0.0, 0.0)):
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs"), P.Range.parse("QQ-JJ,AQs,AJs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"), river=P.card_from_str("3s"))
    tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=100,
                      river_bet_sizes=[(bets, raises), (bets, raises)], rake_rate=rake[0], rake_cap=rake[1])
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def turn_game(oop="AA,KK,AKs", ip="QQ-JJ,AQs", board="Qs Jh 2c", turn="8d", rake=(0.0, 0.0)):
    cc = P.CardConfig(range=(P.Range.parse(oop), P.Range.parse(ip)), flop=P.flop_from_str(board),
                      turn=P.card_from_str(turn))
    s = [("50%, a", ""), ("50%, a", "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=100, effective_stack=200, turn_bet_sizes=s,
                      river_bet_sizes=s, rake_rate=rake[0], rake_cap=rake[1])
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def same(a, b, what):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.shape != b.shape or not np.array_equal(a, b):
        n = int((a != b).sum()) if a.shape == b.shape else -1
        raise AssertionError(f"{what}: {n} entries differ")


def check_stepwise(game, iters, storage="f32"):
    b = P.GpuBatchSolver(0, storage=storage)
    b.prepare_batch([game], iters + 1)
    o = O.OracleGame(game.serialize(), 16 if storage == "compressed" else 32)
    for t in range(iters):
        b.run_iterations(t, t + 1)
        o.solve_step(t)
        if storage == "compressed":
            r, s, rs, ss = b.raw_state(0)
            orr, os_, ors, oss = o.compressed_state()
            same(r.view(np.int16), orr, f"iteration {t} regret words")
            same(s, os_, f"iteration {t} strategy words")
            same(rs, ors, f"iteration {t} regret scales")
            same(ss, oss, f"iteration {t} strategy scales")
        else:
            r, s = o.state()
            same(b.regrets(0), r, f"iteration {t} regrets")
            same(b.strategy_sum(0), s, f"iteration {t} strategy sums")
    e = float(b.exploitability()[0])
    if abs(e - o.exploitability()) > 1e-6 * max(1.0, abs(o.exploitability())):
        raise AssertionError(f"exploitability {e} vs oracle {o.exploitability()}")
    if storage == "f32":
        cf = b.root_cfvalues()[0]
        same(cf[0], o.root_cfvalues(0), "root cfvalues player 0")
        same(cf[1], o.root_cfvalues(1), "root cfvalues player 1")


def run_std_stepwise():
    check_stepwise(std_game(), 6)


def run_std_raises_stepwise():
    check_stepwise(std_game("33%, 75%, a", "2x, a"), 4)


def run_std_solve():
    g = std_game()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 60, [1.0])
    e, it = b.solve(60, 5)
    o = O.OracleGame(g.serialize())
    eo, ito = o.solve(60, 1.0)
    if int(it[0]) != ito or float(e[0]) != np.float32(eo):
        raise AssertionError(f"solve: ({float(e[0])}, {int(it[0])}) vs oracle ({eo}, {ito})")


def run_batch():
    """Three games of different sizes in one batch (work queue order, rows of
    different strides), each compared with its own oracle."""
    games = [std_game(), std_game("33%, a"), W.river_batch(1, seed=5, iterations=10)[0].build()]
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 8)
    b.run_iterations(0, 3)
    for i, g in enumerate(games):
        o = O.OracleGame(g.serialize())
        for t in range(3):
            o.solve_step(t)
        r, s = o.state()
        same(b.regrets(i), r, f"game {i} regrets")
        same(b.strategy_sum(i), s, f"game {i} strategy sums")


def run_turn_stepwise():
    """A turn game: root tile plus one subtree tile per river card, driven by
    the phase kernel; then the FP64 best response through the phase kernel.
    On a device of 64 CUs, so the root tile's forward and backward passes
    and the chance sums and terminals of its leaves run split over three
    workgroups by hand slot."""
    os.environ["EMU_NUM_CUS"] = "64"
    g = turn_game("AA,AKs", "KK,QQ")
    check_stepwise(g, 2)
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 2)
    b.run_iterations(0, 2)
    e64 = float(b.exploitability_f64()[0])
    o = O.OracleGame(g.serialize(), 64)
    o.set_state(b.regrets(0).astype(np.float64), b.strategy_sum(0).astype(np.float64))
    eo = o.exploitability()
    if abs(e64 - eo) > 1e-9 * max(1.0, abs(eo)):
        raise AssertionError(f"turn fp64 exploitability {e64} vs oracle {eo}")


def run_compressed_stepwise():
    check_stepwise(std_game(), 6, storage="compressed")


def run_fp64_exploitability():
    """FP64 best response vs the f64 oracle's best response of the same
    strategy sums, on two river trees."""
    for g in (std_game(), std_game("33%, 75%, a", "2x, a")):
        b = P.GpuBatchSolver(0)
        b.prepare_batch([g], 10)
        b.run_iterations(0, 10)
        e64 = float(b.exploitability_f64()[0])
        o = O.OracleGame(g.serialize(), 64)
        o.set_state(b.regrets(0).astype(np.float64), b.strategy_sum(0).astype(np.float64))
        eo = o.exploitability()
        if abs(e64 - eo) > 1e-9 * max(1.0, abs(eo)):
            raise AssertionError(f"fp64 exploitability {e64} vs oracle {eo}")


def check_fp64(g, iters):
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], iters)
    b.run_iterations(0, iters)
    e64 = float(b.exploitability_f64()[0])
    o = O.OracleGame(g.serialize(), 64)
    o.set_state(b.regrets(0).astype(np.float64), b.strategy_sum(0).astype(np.float64))
    eo = o.exploitability()
    if abs(e64 - eo) > 1e-9 * max(1.0, abs(eo)):
        raise AssertionError(f"fp64 exploitability {e64} vs oracle {eo}")


def run_raked():
    """Raked river (5% capped at 3 chips; the cap binds on the larger pots)
    and raked turn games: raked showdown values, state after every
    iteration, the raked exploitability (best response minus the value of the
    average strategy) in f32 and f64, and the solve schedule."""
    rake = (0.05, 3.0)
    check_stepwise(std_game("50%, a", "", rake), 5)
    check_stepwise(std_game("33%, 75%, a", "2x, a", (0.1, 100.0)), 3)
    g = std_game("50%, a", "", rake)
    check_fp64(g, 8)
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 40, [1.0])
    e, it = b.solve(40, 5)
    eo, ito = O.OracleGame(g.serialize()).solve(40, 1.0)
    if int(it[0]) != ito or float(e[0]) != np.float32(eo):
        raise AssertionError(f"raked solve: ({float(e[0])}, {int(it[0])}) vs oracle ({eo}, {ito})")
    t = turn_game("AA,AKs", "KK,QQ", rake=rake)
    check_stepwise(t, 2)
    check_fp64(t, 2)


def run_reach_at_nodes():
    """batch.rs:765-797 compute_reach_at_nodes on the device (river and turn
    games, every node of each) against the oracle's walk under the average
    strategy of the same state."""
    for g, iters in ((std_game("33%, 75%, a", "2x, a"), 6), (turn_game("AA,AKs", "KK,QQ"), 2)):
        b = P.GpuBatchSolver(0)
        b.prepare_batch([g], iters)
        e, it = b.solve(iters, 0)
        o = O.OracleGame(g.serialize())
        o.set_state(b.regrets(0).astype(np.float64), b.strategy_sum(0).astype(np.float64))
        nodes = np.arange(g.num_nodes())
        for p in range(2):
            got = b.reach_at_nodes(0, p, nodes)
            for n in nodes:
                same(got[n], np.float32(o.reach_at_node(int(n), p)), f"reach of {p} at node {n}")


def flop_game(oop="AA,KK", ip="QQ,AKs", board="Qs Jh 2c", bets="a"):
    cc = P.CardConfig(range=(P.Range.parse(oop), P.Range.parse(ip)), flop=P.flop_from_str(board))
    s = [(bets, ""), (bets, "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Flop, starting_pot=100, effective_stack=100, flop_bet_sizes=s,
                      turn_bet_sizes=s, river_bet_sizes=s)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def run_flop_stepwise():
    """A flop game on a monotone flop with suit-symmetric ranges: root tile
    down to the turn deals (isomorphic turn cards folded in by the swaps),
    one tile per turn card holding its river deals (isomorphic rivers too)
    and river subtrees, so chance nodes inside a tile and terminals of many
    runouts in one tile; state, exploitability and root cfvalues against the
    oracle.  On a device of 64 CUs (the flop tile's passes and leaves split
    by hand slot)."""
    os.environ["EMU_NUM_CUS"] = "64"
    check_stepwise(flop_game("AA,KK", "QQ,AK", "2s 7s Js"), 1)


def run_wave_path():
    """Subtree tiles of games with at most 256 hands per player run one per
    wave (eight tiles per workgroup, each wave with its own LDS region): a
    turn game and a raked turn game with every subtree tile on that path
    (flop_stepwise and turn_stepwise take it too), state after every
    iteration against the oracle and against the level-by-level path, then
    exploitability, root cfvalues and the reach rows download_reach
    rebuilds."""
    games = [turn_game("AA,AKs", "KK,QQ"), turn_game("KK,AQs", "AA,JJ", rake=(0.05, 30.0))]
    runs = []
    for dfs in (True, False):
        b = P.GpuBatchSolver(0)
        b.set_depth_first(dfs)
        b.prepare_batch(games, 4)
        info = b.depth_first_info()
        if dfs and (info["wave_tiles"] == 0 or info["wave_tiles"] != info["tiles"]):
            raise AssertionError(f"wave path tiles {info}")
        if not dfs and info["wave_tiles"] != 0:
            raise AssertionError(f"wave path tiles with the depth-first path off: {info}")
        runs.append(b)
    b, lv = runs
    oracles = [O.OracleGame(g.serialize()) for g in games]
    for t in range(2):
        b.run_iterations(t, t + 1)
        lv.run_iterations(t, t + 1)
        for i, o in enumerate(oracles):
            o.solve_step(t)
            r, s = o.state()
            same(b.regrets(i), r, f"game {i} iteration {t} regrets")
            same(b.strategy_sum(i), s, f"game {i} iteration {t} strategy sums")
            same(b.regrets(i), lv.regrets(i), f"game {i} iteration {t} regrets vs level-by-level")
    same(b.download_reach(64), lv.download_reach(64), "download_reach vs level-by-level")
    e = b.exploitability()
    same(e, lv.exploitability(), "exploitability vs level-by-level")
    for i, o in enumerate(oracles):
        if abs(float(e[i]) - o.exploitability()) > 1e-6 * max(1.0, abs(o.exploitability())):
            raise AssertionError(f"game {i} exploitability {e[i]} vs oracle {o.exploitability()}")
        cf = b.root_cfvalues()[i]
        for p in range(2):
            same(cf[p], o.root_cfvalues(p), f"game {i} root cfvalues player {p}")


def run_half_storage():
    """FP16 / BF16 words with per-node scales: the solve converges like the
    FP64 oracle on the same tree (no reference counterpart)."""
    g = std_game()
    o = O.OracleGame(g.serialize(), 64)
    eo, _ = o.solve(100)
    for storage in ("fp16", "bf16"):
        b = P.GpuBatchSolver(0, storage=storage)
        b.prepare_batch([g], 100)
        e, _ = b.solve(100, 0)
        cf = b.root_cfvalues()[0]
        err = max(np.abs(cf[0] - o.root_cfvalues(0)).max(), np.abs(cf[1] - o.root_cfvalues(1)).max())
        if not (abs(float(e[0]) - eo) <= 0.5 and err <= 1.0):   # chips, pot 100
            raise AssertionError(f"{storage}: expl {float(e[0])} vs fp64 {eo}, root cfv error {err}")


def run_mixed_batch():
    """River games (whole-game kernel) and a turn game (phase launches) in one
    batch: every game's exploitability and iteration count as its own oracle
    solve (the turn loop must not overwrite the river games' results)."""
    games = [std_game(), turn_game("AA", "KK,QQ"), std_game("33%, a")]
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 10)
    e, it = b.solve(10, 5)
    for i, g in enumerate(games):
        o = O.OracleGame(g.serialize())
        eo, ito = o.solve(10)
        if int(it[i]) != ito or float(e[i]) != np.float32(eo):
            raise AssertionError(f"game {i}: ({float(e[i])}, {int(it[i])}) vs oracle ({eo}, {ito})")


def boundary_game(depth=0, flop="Qs Jh 2c", turn="8d", oop="AA,KK,QQ,AKs", ip="QQ-JJ,AQs,AJs"):
    """Turn-rooted tree cut at the river deal (depth_limit 0): every line that
    reaches the river ends in a depth-boundary leaf."""
    cc = P.CardConfig(range=(P.Range.parse(oop), P.Range.parse(ip)), flop=P.flop_from_str(flop),
                      turn=P.card_from_str(turn) if turn else P.NOT_DEALT)
    s = [("50%, a", "2x"), ("50%, a", "2x")]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn if turn else P.BoardState.Flop, starting_pot=100,
                      effective_stack=200, flop_bet_sizes=s, turn_bet_sizes=s, river_bet_sizes=s, depth_limit=depth)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def set_boundaries(g, targets, seed=3):
    """Seeded values on every boundary; odd ordinals raw, even scaled."""
    rng = np.random.default_rng(seed)
    for b in range(g.num_boundary_nodes()):
        for q in range(2):
            v = rng.uniform(-1.0, 1.0, g.num_private_hands(q)).astype(np.float32)
            for t in targets:
                t(b, q, v, b % 2 == 1)


def run_boundary_stepwise():
    g = boundary_game()
    assert g.num_boundary_nodes() > 2
    bt = P.GpuBatchSolver(0)
    bt.prepare_batch([g], 6)
    o = O.OracleGame(g.serialize())
    set_boundaries(g, [lambda b, q, v, raw: bt.set_boundary_cfvs(0, b, q, v, raw),
                       lambda b, q, v, raw: o.set_boundary_cfvs(b, q, v, raw)])
    for t in range(4):
        bt.run_iterations(t, t + 1)
        o.solve_step(t)
        r, st = o.state()
        same(bt.regrets(0), r, f"iteration {t} regrets")
        same(bt.strategy_sum(0), st, f"iteration {t} strategy sums")
        for b in range(g.num_boundary_nodes()):
            for q in range(2):
                same(bt.boundary_reach(0, b, q), o.boundary_reach(b, q), f"iteration {t} boundary {b} reach {q}")
    e = float(bt.exploitability()[0])
    if abs(e - o.exploitability()) > 1e-6 * max(1.0, abs(o.exploitability())):
        raise AssertionError(f"exploitability {e} vs oracle {o.exploitability()}")


def run_depth_first():
    """The depth-first path (rows in LDS): seeded river subgames, a raked
    river game and a depth-limited turn game with boundary leaves, all taking
    that path, equal to the oracle after every iteration and to the
    level-by-level path at the end (state, exploitability, solve() stop)."""
    games = [s.build() for s in W.river_batch(2, seed=11, iterations=10)]
    games += [std_game("33%, 75%, a", "2x, a", rake=(0.05, 30.0)), boundary_game()]
    runs = []
    for dfs in (True, False):
        b = P.GpuBatchSolver(0)
        b.set_depth_first(dfs)
        b.prepare_batch(games, 12, [0.5] * len(games))
        info = b.depth_first_info()
        if info["tiles"] != (len(games) if dfs else 0):
            raise AssertionError(f"depth-first games {info} (enabled: {dfs})")
        set_boundaries(games[-1], [lambda k, q, v, raw: b.set_boundary_cfvs(len(games) - 1, k, q, v, raw)])
        runs.append(b)
    b, lv = runs
    oracles = []
    for g in games:
        o = O.OracleGame(g.serialize())
        if g.num_boundary_nodes():
            set_boundaries(g, [lambda k, q, v, raw: o.set_boundary_cfvs(k, q, v, raw)])
        oracles.append(o)
    for t in range(3):
        b.run_iterations(t, t + 1)
        lv.run_iterations(t, t + 1)
        for i, o in enumerate(oracles):
            o.solve_step(t)
            r, s = o.state()
            same(b.regrets(i), r, f"game {i} iteration {t} regrets")
            same(b.strategy_sum(i), s, f"game {i} iteration {t} strategy sums")
            same(b.regrets(i), lv.regrets(i), f"game {i} iteration {t} regrets vs level-by-level")
    same(b.exploitability(), lv.exploitability(), "exploitability vs level-by-level")
    for i, o in enumerate(oracles):
        e = float(b.exploitability()[i])
        if abs(e - o.exploitability()) > 1e-6 * max(1.0, abs(o.exploitability())):
            raise AssertionError(f"game {i} exploitability {e} vs oracle {o.exploitability()}")
    for x in runs:
        x.reset()
    e1, i1 = b.solve(12, 4)
    e2, i2 = lv.solve(12, 4)
    same(e1, e2, "solve() exploitability vs level-by-level")
    same(i1, i2, "solve() iterations vs level-by-level")
    for i in range(len(games)):
        same(b.strategy_sum(i), lv.strategy_sum(i), f"game {i} solve() strategy sums vs level-by-level")
        for p in range(2):
            same(b.root_cfvalues(compute=False)[i][p], lv.root_cfvalues(compute=False)[i][p],
                 f"game {i} root cfvalues {p} vs level-by-level")


def run_expected_values():
    """query.rs:498-760 at the root: normalized weights, per-action detail and
    expected values of both players, bit for bit with the oracle."""
    for g in (std_game("33%, 75%, a", "2x, a"), turn_game("AA,AKs", "KK,QQ")):
        b = P.GpuBatchSolver(0)
        b.prepare_batch([g], 8)
        b.run_iterations(0, 8)
        o = O.OracleGame(g.serialize())
        for t in range(8):
            o.solve_step(t)
        for p in range(2):
            same(b.normalized_weights(0, p), o.normalized_weights(p), f"normalized weights {p}")
            same(b.expected_values_detail(0, p), o.expected_values(p, detail=True), f"detail {p}")
            same(b.expected_values(0, p), o.expected_values(p), f"expected values {p}")


def bf16(a):
    """Round to BF16 (nearest even) and back, as the device stores operands."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def mlp_case(dims, n, seed, max_batch=64):
    from poker_solver_rust_amd.value_net import ValueNet
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mlp_oracle as M
    rng = np.random.default_rng(seed)
    net = ValueNet(dims, max_batch=max_batch, seed=seed)
    p = net.get_params()
    x = rng.normal(0, 1, (n, dims[0])).astype(np.float32)
    y = rng.normal(0, 1, (n, dims[-1])).astype(np.float32)
    mask = (rng.random((n, dims[-1])) < 0.8).astype(np.float32)
    return net, M, p, x, y, mask


def run_mlp_forward():
    """Forward pass (MFMA GEMM tiles, edges of ragged 64 x 64 tiles, bias +
    GELU epilogue) against the oracle on BF16-rounded weights and inputs."""
    dims = (70, 40, 33, 19)
    net, M, p, x, y, mask = mlp_case(dims, 37, 3)
    got = net.forward(x)
    ref = M.forward(np.concatenate([bf16(a) if k % 2 == 0 else a for k, a in
                                    enumerate(np.split(p, np.cumsum([s for i in range(len(dims) - 1)
                                                                     for s in (dims[i] * dims[i + 1], dims[i + 1])])[:-1]))]),
                    dims, bf16(x).astype(np.float64))
    err = np.abs(got - ref).max() / max(np.abs(ref).max(), 1e-9)
    if err > 2e-2:
        raise AssertionError(f"forward relative error {err}")


def run_mlp_gradients(n=37, max_batch=64, dims=(70, 40, 33, 19)):
    """Forward + backward against the oracle: the weight gradients read both
    operands with rows and columns swapped (X^T dZ), the input gradients
    read the weights transposed; every width is padded to 8 columns."""
    net, M, p, x, y, mask = mlp_case(dims, n, 4, max_batch)
    loss, g = net.gradients(x, y, mask, 1.0)
    lo, go = M.loss_and_grads(p.astype(np.float64), dims, x.astype(np.float64), y.astype(np.float64),
                              mask.astype(np.float64), 1.0)
    if abs(loss - lo) > 2e-2 * abs(lo):
        raise AssertionError(f"loss {loss} vs oracle {lo}")
    err = np.abs(g - go).max() / np.abs(go).max()
    print(f"loss {loss:.6g} oracle {lo:.6g}; max gradient error {err:.3g} of the largest gradient", flush=True)
    if err > 5e-2:
        raise AssertionError(f"gradient relative error {err}")


def run_mlp_gradients_split_k():
    """300 rows of a 300 -> 40 -> 300 -> 19 net: every product with K = 300
    (both forward epilogues, the GELU-gradient product of layer 2, the
    weight gradients over the rows) is split in two K slices of 192 and 108,
    the second ending in a ragged step; the bias gradients are summed over
    three row slices."""
    run_mlp_gradients(300, 512, (300, 40, 300, 19))


def run_mlp_gradients_one_row():
    """A batch of one row: K = 1 for the weight gradients (a single, ragged
    K step), one output row per product."""
    run_mlp_gradients(1, 64)


def run_mlp_adamw_step():
    """One training step on a resident batch: every parameter whose oracle
    gradient is not negligible moves by the oracle's AdamW update (at step 1
    the update is lr * g / (|g| + eps), weights decayed)."""
    dims = (70, 40, 33, 19)
    net, M, p, x, y, mask = mlp_case(dims, 32, 5)
    net.set_dataset(x, y, mask)
    net.train(1, 32, lr=1e-2, weight_decay=0.01, delta=1.0, seed=0)
    p1 = net.get_params().astype(np.float64)
    _, go = M.loss_and_grads(p.astype(np.float64), dims, x.astype(np.float64), y.astype(np.float64),
                             mask.astype(np.float64), 1.0)
    ref = M.AdamW(dims, 1e-2, 0.01).step(p.astype(np.float64), go)
    big = np.abs(go) > 1e-2 * np.abs(go).max()
    err = np.abs(p1 - ref)[big].max()
    print(f"AdamW step: {int(big.sum())} of {len(p)} parameters compared, max deviation {err:.3g} (lr 1e-2)", flush=True)
    if err > 1e-3:
        raise AssertionError(f"AdamW step deviates by {err}")


def run_mlp_adamw_clipped():
    """One training step with gradient-norm clipping: max_norm puts the
    median clipped gradient at AdamW's eps, so the step is lr * g_c / (|g_c|
    + eps) with a coefficient-dependent size.  The device step matches the
    oracle's clipped step, misses its unclipped step, and the reported norm
    is the oracle's gradient norm."""
    dims = (70, 40, 33, 19)
    net, M, p, x, y, mask = mlp_case(dims, 32, 6)
    p = p.astype(np.float64)
    _, go = M.loss_and_grads(p, dims, x.astype(np.float64), y.astype(np.float64), mask.astype(np.float64), 1.0)
    gc, norm = M.clip_grad_norm(go, 1e-8 * np.linalg.norm(go) / np.median(np.abs(go)))
    net.set_dataset(x, y, mask)
    net.set_grad_clip(1e-8 * np.linalg.norm(go) / np.median(np.abs(go)))
    net.train(1, 32, lr=1e-2, weight_decay=0.01, delta=1.0, seed=0)
    d = net.get_params().astype(np.float64)
    big = np.abs(go) > 1e-2 * np.abs(go).max()
    err = np.abs(d - M.AdamW(dims, 1e-2, 0.01).step(p, gc))[big].max()
    err_plain = np.abs(d - M.AdamW(dims, 1e-2, 0.01).step(p, go))[big].max()
    dn = net.last_grad_norm()
    print(f"clipped AdamW step: max deviation {err:.3g} from the clipped oracle step, {err_plain:.3g} from the "
          f"unclipped one; norm {dn:.6g} vs {norm:.6g}", flush=True)
    if err > 1e-3 or err_plain < 3e-3 or abs(dn - norm) > 2e-2 * norm:
        raise AssertionError(f"clipped AdamW step: {err}, {err_plain}, {dn} vs {norm}")


CHECKS = {name[4:]: fn for name, fn in globals().items() if name.startswith("run_")}

if __name__ == "__main__":
    import time
    for name in sys.argv[1:]:
        print(f"check {name}", flush=True)
        t0 = time.time()
        CHECKS[name]()
        print(f"check {name} passed in {time.time() - t0:.1f} s", flush=True)
    sys.exit(0)
