"""Expected values at the root (query.rs:498-760) in the oracle, pinned
without the reference:

  - normalized weights against the brute-force definition: w_i times the
    weight of the opponent hands that share no card with hand i (dead hands
    zeroed, query.rs:886-920);
  - the pot identity: under any strategy profile the cfvalues are zero-sum
    (unraked), so sum_i nw0_i EV0_i + sum_j nw1_j EV1_j = pot * sum nw (both
    sums of normalized weights are the joint weight of compatible pairs);
  - a complete-information game (one hand each, the winner is fixed): after
    solving, the winner's chip EV is the whole pot and the loser's 0;
  - the acting player's EV is the average-strategy mix of its per-action
    detail.
The device values are compared with the oracle in test_kernel_emu.py and
test_gpu_parity.py.
"""
import numpy as np
import pytest

import oracle_binding as O
import poker_solver_rust_amd as P
from poker_solver_rust_amd import workloads as W


def std_game(bets="50%, a", raises=""):
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs"), P.Range.parse("QQ-JJ,AQs,AJs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"), river=P.card_from_str("3s"))
    tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=100,
                      river_bet_sizes=[(bets, raises), (bets, raises)])
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def turn_game():
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,AKs"), P.Range.parse("QQ-JJ,AQs")), flop=P.flop_from_str("Qs Jh 2c"),
                      turn=P.card_from_str("8d"))
    s = [("50%, a", ""), ("50%, a", "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=60, effective_stack=200, turn_bet_sizes=s,
                      river_bet_sizes=s)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def test_normalized_weights_brute_force():
    g = std_game()
    o = O.OracleGame(g.serialize())
    for p in range(2):
        nw = o.normalized_weights(p)
        pc, oc = g.private_cards(p), g.private_cards(p ^ 1)
        w, wo = g.initial_weights(p), g.initial_weights(p ^ 1)
        ref = np.array([w[i] * sum(wo[j] for j, (c, d) in enumerate(oc) if not ({a, b} & {c, d}))
                        for i, (a, b) in enumerate(pc)])
        assert np.allclose(nw, ref, rtol=1e-6)


@pytest.mark.parametrize("make,iters", [(std_game, 200), (turn_game, 30), (lambda: W.river_uniform(5).build(), 60)])
def test_pot_identity(make, iters):
    g = make()
    o = O.OracleGame(g.serialize())
    for t in range(iters):
        o.solve_step(t)
    nw = [o.normalized_weights(p).astype(np.float64) for p in range(2)]
    ev = [o.expected_values(p).astype(np.float64) for p in range(2)]
    pot = g.serialize()  # starting pot is the first i32 after magic + version
    starting_pot = int.from_bytes(pot[8:12], "little", signed=True)
    total = float(nw[0] @ ev[0] + nw[1] @ ev[1])
    assert abs(nw[0].sum() - nw[1].sum()) <= 1e-6 * nw[0].sum()
    assert abs(total / nw[0].sum() - starting_pot) <= 1e-4 * starting_pot


def test_complete_information_expected_values():
    a = np.zeros(1326, np.float32)
    b = np.zeros(1326, np.float32)
    a[P.card_pair_to_index(P.card_from_str("Ah"), P.card_from_str("Ad"))] = 1.0
    b[P.card_pair_to_index(P.card_from_str("Kh"), P.card_from_str("Kd"))] = 1.0
    cc = P.CardConfig(range=(P.Range(a), P.Range(b)), flop=P.flop_from_str("2c7s9h"), turn=P.card_from_str("Js"),
                      river=P.card_from_str("3c"))
    tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=100,
                      river_bet_sizes=[("50%, a", "2x"), ("50%, a", "2x")])
    g = P.PostFlopGame.with_config(cc, P.ActionTree(tc))
    o = O.OracleGame(g.serialize())
    o.solve(300)
    assert abs(o.expected_values(0)[0] - 100.0) < 0.1
    assert abs(o.expected_values(1)[0] - 0.0) < 0.1


def test_acting_player_mixes_detail_with_average_strategy():
    g = std_game("33%, 75%, a")
    o = O.OracleGame(g.serialize())
    for t in range(50):
        o.solve_step(t)
    h = g.num_private_hands(0)
    det = o.expected_values(0, detail=True)
    a = len(det) // h
    assert a > 1 and len(det) == a * h
    s = o.state()[1][:a * h].astype(np.float32).reshape(a, h)
    d = s.sum(0)
    sig = np.where(d == 0, 1.0 / a, s / np.where(d == 0, 1, d))
    mix = (det.reshape(a, h).astype(np.float64) * sig).sum(0)
    assert np.allclose(o.expected_values(0), mix, rtol=1e-5, atol=1e-4)
    assert np.array_equal(o.expected_values(1, detail=True), o.expected_values(1))
