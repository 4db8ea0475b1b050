"""Hand evaluator, hand table and game build against the reference's fixtures:
hand_table.rs (4824 sorted encodings, committed as tests/golden/
hand_table_ref.txt), hand.rs:440-448 category counts, examples/basic.rs:51,121
hand lists, and the standard river fixture (solver.rs:778-798)."""
import os
import re
import subprocess

import numpy as np
import pytest

import poker_solver_rust_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "poker_solver_rust_amd", "csrc")


def test_hand_table_matches_reference_fixture():
    ref = [int(x) for x in open(os.path.join(ROOT, "tests", "golden", "hand_table_ref.txt"))]
    text = open(os.path.join(CSRC, "hand_table.inc")).read().split("{", 1)[1]
    mine = [int(v) for v in re.findall(r"-?\d+", text)]
    assert len(ref) == 4824 and mine == ref


@pytest.mark.slow
def test_hand_table_generator_and_category_counts(tmp_path):
    """Enumerates all C(52,7) hands: the table it writes must be the committed
    one and the category counts the textbook ones (hand.rs:440-448)."""
    exe = tmp_path / "gen"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-o", str(exe), os.path.join(CSRC, "gen_hand_table.cpp")])
    out = tmp_path / "t.inc"
    counts = subprocess.check_output([str(exe), str(out), "--counts"]).decode().split()
    assert [int(c) for c in counts] == [23294460, 58627800, 31433400, 6461620, 6180020, 4047644, 3473184, 224848,
                                        41584]
    assert out.read_text() == open(os.path.join(CSRC, "hand_table.inc")).read()


def cards(s):
    return [P.card_from_str(s[i:i + 2]) for i in range(0, len(s), 2)]


def test_evaluator_orders_categories():
    # royal flush > quads > full house > flush > straight > trips > two pair > pair > high card
    hands = ["AsKsQsJsTs2c3d", "AsAhAdAc2s3c4d", "AsAhAdKcKs2c3d", "As9s7s5s3s2c4d", "AsKdQcJhTs2c3d",
             "AsAhAd9c7s2c3d", "AsAhKdKc7s2c3d", "AsAh9d8c7s2c3d", "AsKh9d8c6s2c3d"]
    s = [P.hand_evaluate(cards(h)) for h in hands]
    assert s == sorted(s, reverse=True)
    # the wheel is the lowest straight, 6-high straight beats it
    assert P.hand_evaluate(cards("As2d3c4h5sKdQc")) < P.hand_evaluate(cards("2d3c4h5s6sKdQc"))
    # kicker beyond the best five does not count
    assert P.hand_evaluate(cards("AsAh9d8c7s3c2d")) == P.hand_evaluate(cards("AsAh9d8c7s4c2d"))
    assert P.hand_evaluate(cards("2s3s4s5s6sAsKs")) > P.hand_evaluate(cards("AsKsQsJs9s2c3d"))


def std_game(bets="50%, a"):
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs"), P.Range.parse("QQ-JJ,AQs,AJs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"), river=P.card_from_str("3s"))
    tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=100,
                      river_bet_sizes=[(bets, ""), (bets, "")])
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def test_standard_river_game():
    g = std_game()
    assert g.num_private_hands(0) == 19 and g.num_private_hands(1) == 12
    # num_combinations: non-conflicting (oop, ip) pairs, interpreter.rs:619-640
    oc, ic = g.private_cards(0), g.private_cards(1)
    n = sum(1 for a in oc for b in ic if not ({a[0], a[1]} & {b[0], b[1]}))
    assert g.num_combinations() == n
    assert oc == sorted(oc)


def hole_str(c):
    a, b = max(c), min(c)
    R, S = "23456789TJQKA", "cdhs"
    return R[a >> 2] + S[a & 3] + R[b >> 2] + S[b & 3]


def test_basic_example_hands():
    """examples/basic.rs:10-51,121: first OOP hands and 250 IP hands."""
    oop = P.Range.parse("66+,A8s+,A5s-A4s,AJo+,K9s+,KQo,QTs+,JTs,96s+,85s+,75s+,65s,54s")
    ip = P.Range.parse("QQ-22,AQs-A2s,ATo+,K5s+,KJo+,Q8s+,J8s+,T7s+,96s+,86s+,75s+,64s+,53s+")
    cc = P.CardConfig(range=(oop, ip), flop=P.flop_from_str("Td9d6h"), turn=P.card_from_str("Qc"))
    s = [("60%, e, a", "2.5x"), ("60%, e, a", "2.5x")]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=200, effective_stack=900, flop_bet_sizes=s,
                      turn_bet_sizes=s, river_bet_sizes=s, river_donk_sizes="50%", add_allin_threshold=1.5,
                      force_allin_threshold=0.15, merging_threshold=0.1)
    g = P.PostFlopGame.with_config(cc, P.ActionTree(tc))
    assert [hole_str(c) for c in g.private_cards(0)[:10]] == [
        "5c4c", "Ac4c", "5d4d", "Ad4d", "5h4h", "Ah4h", "5s4s", "As4s", "6c5c", "7c5c"]
    assert g.num_private_hands(1) == 250


def test_game_config_errors():
    tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=100,
                      river_bet_sizes=[("50%", ""), ("50%", "")])
    tree = P.ActionTree(tc)
    full = P.Range.ones()
    for cc in [P.CardConfig(range=(P.Range(), full), flop=[0, 1, 2], turn=3, river=4),     # empty OOP range
               P.CardConfig(range=(full, full), flop=[0, 0, 2], turn=3, river=4),          # duplicate flop
               P.CardConfig(range=(full, full), flop=[0, 1, 2], turn=3),                   # state mismatch
               P.CardConfig(range=(full, full), flop=[0, 1, 2], turn=3, river=3)]:        # river == turn
        with pytest.raises(P.PsrError):
            P.PostFlopGame.with_config(cc, tree)
    # the only combos share a card with each other: no valid assignment
    a = np.zeros(1326, np.float32)
    a[P.card_pair_to_index(40, 41)] = 1.0
    b = np.zeros(1326, np.float32)
    b[P.card_pair_to_index(40, 42)] = 1.0
    with pytest.raises(P.PsrError):
        P.PostFlopGame.with_config(P.CardConfig(range=(P.Range(a), P.Range(b)), flop=[0, 1, 2], turn=3, river=4), tree)


def test_turn_game_has_48_rivers():
    from poker_solver_rust_amd import workloads as W
    g = W.turn_uniform01(0).build()
    assert g.num_private_hands(0) == 1128
    assert g.num_nodes() > 10000


def test_topology_parents_and_depths():
    """psr_game_topology: breadth-first arena, children contiguous, depth =
    parent depth + 1, so edge e of the reference topology leads to node
    e + 1 and parents are non-decreasing along the arena."""
    cc = P.CardConfig(range=(P.Range.parse("AA,KK"), P.Range.parse("QQ,AKs")), flop=P.flop_from_str("Qs Jh 2c"),
                      turn=P.card_from_str("8d"))
    s = [("50%, a", ""), ("50%, a", "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=100, effective_stack=200, turn_bet_sizes=s,
                      river_bet_sizes=s)
    g = P.PostFlopGame.with_config(cc, P.ActionTree(tc))
    par, dep = g.topology()
    assert par[0] == -1 and dep[0] == 0
    assert np.all(np.diff(par[1:]) >= 0)
    assert np.all(dep[1:] == dep[par[1:]] + 1)
    assert dep.max() >= 4   # turn betting, chance, river betting
