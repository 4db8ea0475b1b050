"""16-bit storage on the GPU.

compressed: the reference's i16 regrets / u16 strategy sums with per-node
scales (allocate_memory(true); utility.rs:142-202, solver.rs:486-553) --
compared word for word with the oracle's compressed mode.
fp16 / bf16: value/scale words with a per-node f32 scale and f32 arithmetic
-- no reference counterpart, so they are measured against the FP64 oracle on
the same tree: exploitability gap and per-hand CFV error in chips.
"""
import numpy as np
import pytest

import oracle_binding as O
import poker_solver_rust_amd as P
from poker_solver_rust_amd import workloads as W

pytestmark = pytest.mark.gpu


def std_game():
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs"), P.Range.parse("QQ-JJ,AQs,AJs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"), river=P.card_from_str("3s"))
    tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=100,
                      river_bet_sizes=[("50%, a", "2x"), ("50%, a", "2x")])
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def turn_game():
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs"), P.Range.parse("QQ-JJ,AQs,AJs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"))
    s = [("50%, a", ""), ("50%, a", "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=100, effective_stack=200, turn_bet_sizes=s,
                      river_bet_sizes=s)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def flop_game():
    cc = P.CardConfig(range=(P.Range.parse("AA,KK"), P.Range.parse("QQ,AKs")), flop=P.flop_from_str("Qs Jh 2c"))
    s = [("50%, a", ""), ("50%, a", "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Flop, starting_pot=100, effective_stack=100, flop_bet_sizes=s,
                      turn_bet_sizes=s, river_bet_sizes=s)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


@pytest.mark.parametrize("make", [std_game, turn_game, lambda: W.river_uniform(1).build(), flop_game])
def test_compressed_words_match_oracle(make):
    """Word for word against the oracle's compressed mode after 12
    iterations; the flop game runs every phase (root, turn and river tiles)
    through the 16-bit update."""
    g = make()
    b = P.GpuBatchSolver(0, storage="compressed")
    b.prepare_batch([g], 40)
    o = O.OracleGame(g.serialize(), 16)
    for t in range(12):
        b.run_iterations(t, t + 1)
        o.solve_step(t)
    r, s, rs, ss = b.raw_state(0)
    orr, os_, ors, oss = o.compressed_state()
    assert np.array_equal(rs, ors) and np.array_equal(ss, oss)
    assert np.array_equal(r.view(np.int16), orr)
    assert np.array_equal(s, os_)
    assert abs(float(b.exploitability()[0]) - o.exploitability()) <= 1e-4 * max(1.0, o.exploitability())


def test_compressed_solve_schedule():
    g = W.river_uniform(2).build()
    b = P.GpuBatchSolver(0, storage="compressed")
    b.prepare_batch([g], 200, [0.3])
    e, it = b.solve(200, 5)
    o = O.OracleGame(g.serialize(), 16)
    eo, ito = o.solve(200, 0.3)
    assert int(it[0]) == ito and abs(float(e[0]) - eo) <= 1e-4 * max(1.0, eo)


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_half_storage_against_fp64_oracle(storage):
    """Reduced-precision solve vs the all-f64 oracle on the same tree: the
    exploitability reached and the root CFVs stay within a small fraction of
    the pot."""
    s = W.river_uniform(3)
    g = s.build()
    b = P.GpuBatchSolver(0, storage=storage)
    b.prepare_batch([g], 300)
    e, it = b.solve(300, 0)
    o = O.OracleGame(g.serialize(), 64)
    eo, _ = o.solve(300)
    cf = b.root_cfvalues()[0]
    gap = float(e[0]) - eo
    cfv_err = max(np.abs(cf[0] - o.root_cfvalues(0)).max(), np.abs(cf[1] - o.root_cfvalues(1)).max())
    print(f"{storage}: expl {float(e[0]):.5f} vs fp64 {eo:.5f} (gap {gap:+.5f} chips), root cfv max abs err {cfv_err:.5f}")
    assert float(e[0]) <= 0.01 * s.pot
    assert abs(gap) <= 0.005 * s.pot
    assert cfv_err <= 0.01 * s.pot


def test_storage_mode_must_precede_prepare():
    b = P.GpuBatchSolver(0)
    b.prepare_batch([std_game()], 5)
    with pytest.raises(P.PsrError):
        P.range_solver._check(P.load_library().psr_batch_set_storage(b._h, 1))


@pytest.mark.parametrize("make", [std_game, turn_game, lambda: W.river_uniform(4).build()])
def test_fp64_exploitability_matches_oracle(make):
    """FP64 best response on the device vs the f64 oracle's best response of
    the very same strategy sums (loaded into an f64 handle)."""
    g = make()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 60)
    b.run_iterations(0, 60)
    e64 = float(b.exploitability_f64()[0])
    o = O.OracleGame(g.serialize(), 64)
    o.set_state(b.regrets(0).astype(np.float64), b.strategy_sum(0).astype(np.float64))
    eo = o.exploitability()
    assert abs(e64 - eo) <= 1e-9 * max(1.0, abs(eo)), (e64, eo)
    # and the f32 pass agrees to f32 precision
    e32 = float(b.exploitability()[0])
    assert abs(e32 - eo) <= 1e-4 * max(1.0, abs(eo))


@pytest.fixture(scope="module")
def fp64_turn_oracle():
    """The FP64 oracle's solve of the seeded turn subgame (turn_uniform01:
    U[0,1] ranges, 48 river cards, ~17k nodes) for C4_ITERS iterations, the
    turn's children on the host's cores as the reference's rayon policy."""
    import os
    s = W.turn_uniform01(20260)
    g = s.build()
    o = O.OracleGame(g.serialize(), 64)
    o.set_threads(max(1, min(16, os.cpu_count() or 1)))
    for t in range(C4_ITERS):
        o.solve_step(t)
    return s, g, o.exploitability(), [o.root_cfvalues(p) for p in range(2)]


C4_ITERS = 100


# measured on the device (100 iterations, pot 60): gaps +0.0006 / +0.0081 /
# +0.0081 chips, root cfv errors 0.16% / 0.22% / 0.12% of the largest CFV
@pytest.mark.parametrize("storage,gap_pot,cfv_rel", [("f32", 0.0005, 0.01), ("fp16", 0.0005, 0.01), ("bf16", 0.0005, 0.01)])
def test_reduced_precision_turn_solve_against_fp64_oracle(fp64_turn_oracle, storage, gap_pot, cfv_rel):
    """Config C4 on its own workload: the turn tree of C2 with regrets and
    strategy sums in FP16 / BF16 words with per-node FP32 scales (f32 for
    reference), C4_ITERS DCFR iterations on the device, then the device's
    FP64 best response (utility.rs:380-392 in f64).  The exploitability gap
    to the FP64 oracle's own solve of the same tree (gap_pot: fraction of the
    pot) and the per-hand root CFV error (cfv_rel: fraction of the largest
    oracle CFV) stay within the stated bounds."""
    s, g, eo, ref = fp64_turn_oracle
    b = P.GpuBatchSolver(0, storage=storage)
    b.prepare_batch([g], C4_ITERS)
    b.run_iterations(0, C4_ITERS)
    e64 = float(b.exploitability_f64()[0])
    cf = b.root_cfvalues()[0]
    scale = max(float(np.abs(r).max()) for r in ref)
    err = max(float(np.abs(cf[p].astype(np.float64) - ref[p]).max()) for p in range(2))
    print(f"C4 {storage}: fp64 exploitability {e64:.6f} vs oracle {eo:.6f} (gap {e64 - eo:+.6f} chips, pot {s.pot}), "
          f"root cfv max abs err {err:.3e} (max |cfv| {scale:.3e})")
    assert e64 >= 0.0
    assert abs(e64 - eo) <= gap_pot * s.pot
    assert err <= cfv_rel * scale
