"""Pins for the CPU oracle (oracle/dcfr_oracle.cpp), independent of the
oracle's own formulas:
  - DiscountParams known values (solver.rs:688-722)
  - fold / showdown values against an O(H^2) pairwise brute force (the
    dense formulation of gpu-range-solver extract.rs:306-341) built from the
    7-card evaluator on random reaches
  - convergence thresholds of the reference's tests (solver.rs:772-805,
    gpu-range-solver lib.rs:214-245)
  - invariants: exploitability >= 0, best response >= current value,
    unraked zero sum
  - a closed-form game value (complete information: the winner takes the pot)
"""
import numpy as np
import pytest

import oracle_binding as O
import poker_solver_rust_amd as P
from poker_solver_rust_amd import workloads as W


def std_game(bets="50%, a", rake=(0.0, 0.0)):
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs"), P.Range.parse("QQ-JJ,AQs,AJs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"), river=P.card_from_str("3s"))
    tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=100,
                      river_bet_sizes=[(bets, ""), (bets, "")], rake_rate=rake[0], rake_cap=rake[1])
    return cc, P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def test_discount_params_known_values():
    a, b, g = O.discount(0)
    assert (a, b, g) == (0.0, 0.5, 0.0)
    a, b, g = O.discount(1)
    assert a == 0.0 and b == 0.5
    assert O.discount(4)[2] == 0.0          # strategy-sum restart at powers of 4
    assert abs(O.discount(5)[2] - 0.125) < 1e-6
    assert O.discount(16)[2] == 0.0 and O.discount(64)[2] == 0.0
    assert O.discount(100)[0] > 0.9 and O.discount(100)[2] > 0.0
    # closed forms: alpha = (t-1)^1.5 / ((t-1)^1.5 + 1); gamma = (t'/(t'+1))^3
    for t in [2, 3, 7, 33, 250, 999]:
        x = (t - 1) ** 1.5
        p4 = 4 ** int(np.floor(np.log(t) / np.log(4)))
        tg = t - p4
        assert O.discount(t)[0] == np.float32(x / (x + 1))
        assert abs(O.discount(t)[2] - (tg / (tg + 1)) ** 3) < 1e-7


def terminal_nodes(blob):
    """(node index, is_fold, folded player, amount) of every terminal node,
    read from the serialized arena (format in game.cpp)."""
    import struct
    off = 8 + 4 + 4 + 8 + 8 + 8 + 8
    hands = []
    for p in range(2):
        (h,) = struct.unpack_from("<I", blob, off)
        off += 4 + 2 * h + 4 * h + 2 * h
        hands.append(h)
    (n,) = struct.unpack_from("<I", blob, off)
    off += 4
    out = []
    for i in range(n):
        player, turn, river, _, amount, first, num, runout = struct.unpack_from("<BBBBiIIi", blob, off)
        off += 20
        if player & 8:
            out.append((i, (player & 24) == 24, player & 3, amount))
    return out


def brute_force(game, board, node, player, reach, pot, rake=(0.0, 0.0)):
    """O(H^2): sum over non-conflicting opponent hands of reach * payoff.
    With rake, the house takes min(pot * rate, cap) from the pot: the winner
    of a fold or showdown nets half the pot minus the rake, and a tie splits
    the raked pot, so each side nets -rake / 2."""
    pc, oc = game.private_cards(player), game.private_cards(player ^ 1)
    ncomb = game.num_combinations()
    half = 0.5 * (pot)
    rk = min(pot * rake[0], rake[1])
    win, lose, tie = (half - rk) / ncomb, -half / ncomb, -0.5 * rk / ncomb
    if node[1]:
        pay = win if node[2] != player else lose
    s_p = [P.hand_evaluate(board + [a, b]) for a, b in pc]
    s_o = [P.hand_evaluate(board + [a, b]) for a, b in oc]
    out = np.zeros(len(pc))
    for i, (a, b) in enumerate(pc):
        acc = 0.0
        for j, (c, d) in enumerate(oc):
            if {a, b} & {c, d}:
                continue
            if node[1]:
                acc += pay * reach[j]
            elif s_p[i] > s_o[j]:
                acc += win * reach[j]
            elif s_p[i] < s_o[j]:
                acc += lose * reach[j]
            else:
                acc += tie * reach[j]
        out[i] = acc
    return out


RAKES = {"std": (0.0, 0.0), "uniform": (0.0, 0.0), "std_raked": (0.05, 3.0), "std_raked_uncapped": (0.1, 1e9)}


@pytest.mark.parametrize("which", list(RAKES))
def test_terminal_values_match_brute_force(which):
    rng = np.random.default_rng(7)
    rake = RAKES[which]
    if which.startswith("std"):
        cc, g = std_game(rake=rake)
        board = list(cc.flop) + [cc.turn, cc.river]
        pot0 = 100
    else:
        s = W.river_uniform(3)
        g = s.build()
        board = s.board
        pot0 = s.pot
    blob = g.serialize()
    og = O.OracleGame(blob)
    terms = terminal_nodes(blob)
    assert any(t[1] for t in terms) and any(not t[1] for t in terms)
    for t in terms[:6] if which == "uniform" else terms:
        for p in range(2):
            reach = rng.random(g.num_private_hands(p ^ 1)).astype(np.float32).astype(np.float64)
            reach[rng.random(len(reach)) < 0.2] = 0.0
            got = og.evaluate(t[0], p, reach)
            ref = brute_force(g, board, t, p, reach, pot0 + 2 * t[3], rake)
            scale = np.abs(ref).max()
            assert np.abs(got - ref).max() <= 1e-6 * scale + 1e-12, (t, p)


def test_convergence_thresholds():
    """solver.rs:772-805: expl < 10 after 100 iterations; gpu lib.rs:214-245:
    below 1.0 (reference GPU) after 500, CPU reaches < 0.01."""
    _, g = std_game()
    og = O.OracleGame(g.serialize())
    e, it = og.solve(100)
    assert e < 10.0 and it == 100
    og = O.OracleGame(g.serialize())
    e, it = og.solve(500)
    assert 0.0 <= e < 0.01


def test_early_termination():
    _, g = std_game()
    og = O.OracleGame(g.serialize())
    e, it = og.solve(1000, target=1.0)
    assert e <= 1.0 and it < 1000 and it % 5 == 0


@pytest.mark.parametrize("rake", [(0.0, 0.0), (0.05, 3.0)])
def test_invariants_along_the_solve(rake):
    _, g = std_game("33%, 75%, a", rake)
    og = O.OracleGame(g.serialize())
    for t in range(60):
        og.solve_step(t)
        if t % 7 == 0:
            mes, cur = og.mes_ev(), og.current_ev()
            e = og.exploitability()
            assert e >= -1e-5
            assert np.all(mes >= cur - 1e-4)
            if rake[0] == 0.0:
                assert abs(cur[0] + cur[1]) <= 1e-4 * max(1.0, np.abs(cur).max())
                assert abs(e - 0.5 * (mes[0] + mes[1])) <= 1e-5 * max(1.0, abs(e))
            else:
                # the house takes chips: the two values sum to minus the
                # expected rake, and exploitability is measured from them
                assert cur[0] + cur[1] < 0.0
                assert abs(e - 0.5 * ((mes[0] - cur[0]) + (mes[1] - cur[1]))) <= 1e-5 * max(1.0, abs(e))


def test_complete_information_value():
    """One hand each: the winner takes half the pot whatever is played, so the
    solved root value is exactly +-half pot per combination."""
    a = np.zeros(1326, np.float32)
    b = np.zeros(1326, np.float32)
    a[P.card_pair_to_index(P.card_from_str("Ah"), P.card_from_str("Ad"))] = 1.0
    b[P.card_pair_to_index(P.card_from_str("Kh"), P.card_from_str("Kd"))] = 1.0
    cc = P.CardConfig(range=(P.Range(a), P.Range(b)), flop=P.flop_from_str("2c7s9h"), turn=P.card_from_str("Js"),
                      river=P.card_from_str("3c"))
    tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=100,
                      river_bet_sizes=[("50%, a", "2x"), ("50%, a", "2x")])
    g = P.PostFlopGame.with_config(cc, P.ActionTree(tc))
    og = O.OracleGame(g.serialize())
    e, _ = og.solve(300)
    assert e < 1e-3
    assert abs(og.root_cfvalues(0)[0] - 50.0) < 0.05
    assert abs(og.root_cfvalues(1)[0] + 50.0) < 0.05


def test_f64_oracle_close_to_f32():
    _, g = std_game()
    o32 = O.OracleGame(g.serialize(), 32)
    o64 = O.OracleGame(g.serialize(), 64)
    e32, _ = o32.solve(200)
    e64, _ = o64.solve(200)
    assert abs(e32 - e64) < 1e-3
    assert np.abs(o32.root_cfvalues(0) - o64.root_cfvalues(0)).max() < 1e-3


def test_encode_known_answers():
    """utility.rs:1063-1081."""
    sc, out = O.encode_signed([1.0, -1.0, 0.5, -0.5])
    assert sc == 1.0 and out[0] == 32767 and out[1] == -32767 and out[2] == 16384 and out[3] == -16384
    sc, out = O.encode_unsigned([1.0, 0.0, 0.5])
    assert sc == 1.0 and out[0] == 65535 and out[1] == 0 and out[2] == 32768
    sc, out = O.encode_signed([0.0, 0.0])
    assert sc == 0.0 and not out.any()


def test_compressed_oracle_tracks_uncompressed():
    """The compressed mode (allocate_memory(true)) converges like f32 and its
    decoded state stays within the 16-bit quantisation of the f32 state."""
    _, g = std_game()
    o32 = O.OracleGame(g.serialize(), 32)
    o16 = O.OracleGame(g.serialize(), 16)
    e32, _ = o32.solve(300)
    e16, _ = o16.solve(300)
    assert e16 < 10.0 and abs(e16 - e32) < 0.05
    s32 = o32.state()[1]
    s16 = o16.state()[1]
    assert np.abs(s16 - s32).max() <= 0.05 * np.abs(s32).max()


def test_parallel_children_give_identical_state():
    """The oracle's parallel traversal of the nodes before the river deal
    (the reference's rayon policy, utility.rs:103-115) leaves every regret
    and strategy sum bit-identical to the sequential one."""
    import poker_solver_rust_amd as P

    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs,AQo"), P.Range.parse("QQ-99,AJs,KQs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"))
    s = [("50%, a", ""), ("50%, a", "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=100, effective_stack=200, turn_bet_sizes=s,
                      river_bet_sizes=s)
    g = P.PostFlopGame.with_config(cc, P.ActionTree(tc))
    a, b = O.OracleGame(g.serialize()), O.OracleGame(g.serialize())
    b.set_threads(4)
    for t in range(6):
        a.solve_step(t)
        b.solve_step(t)
    ra, sa = a.state()
    rb, sb = b.state()
    assert np.array_equal(ra, rb) and np.array_equal(sa, sb)
    assert a.exploitability() == b.exploitability()
    assert np.array_equal(a.root_cfvalues(0), b.root_cfvalues(0))
