"""Bet-size grammar and action-tree construction against the reference's
tests (crates/range-solver/src/bet_size.rs:244-372,
action_tree.rs:1542-2160, examples/basic.rs:104-122)."""
import pytest

import poker_solver_rust_amd as P
from poker_solver_rust_amd import Action as A

GOOD = ["0%", "75%", "112.5%", "1.001x", "3.5X", "0c", "123C", "0c1r", "100C100R", "e", "E", "2e", "E37.5%",
        "100e.5%", "a", "A"]
BAD = ["0", "1.23", "%", "+42%", "-30%", "x", "0x", "1x", "c", "12.3c", "10c10", "42cr", "c3r", "0c0r",
       "123c101r", "1c2r3", "12c3.4r", "0e", "2.7e", "101e", "3e7", "E%", "1e2e3", "bet", "1a", "a1"]


def river_tree(bet, raise_, pot, stack, add=1.5, force=0.15, merge=0.12, **kw):
    cfg = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=pot, effective_stack=stack,
                       river_bet_sizes=[(bet, raise_), (bet, raise_)], add_allin_threshold=add,
                       force_allin_threshold=force, merging_threshold=merge, **kw)
    return P.ActionTree(cfg)


@pytest.mark.parametrize("s", GOOD)
def test_bet_size_grammar_accepts(s):
    river_tree("", s, 100, 100)


@pytest.mark.parametrize("s", BAD)
def test_bet_size_grammar_rejects(s):
    with pytest.raises(P.PsrError):
        river_tree("", s, 100, 100)


def test_bet_sizes_not_allowed_for_bets():
    with pytest.raises(P.PsrError):
        river_tree("2.5x", "", 100, 100)
    with pytest.raises(P.PsrError):
        river_tree(",", "", 100, 100)
    with pytest.raises(P.PsrError):
        river_tree("10c2r", "", 100, 100)


def test_config_rejections():
    for kw in [dict(starting_pot=0, effective_stack=100), dict(starting_pot=100, effective_stack=0),
               dict(starting_pot=100, effective_stack=100, rake_rate=-0.1),
               dict(starting_pot=100, effective_stack=100, rake_rate=1.1),
               dict(starting_pot=100, effective_stack=100, rake_cap=-1.0),
               dict(starting_pot=100, effective_stack=100, add_allin_threshold=-1.0),
               dict(starting_pot=100, effective_stack=100, force_allin_threshold=-1.0),
               dict(starting_pot=100, effective_stack=100, merging_threshold=-0.1)]:
        with pytest.raises(P.PsrError):
            P.ActionTree(P.TreeConfig(**kw))


def acts(tree, hist=()):
    return tree.available_actions([(int(k), a) for k, a in hist])


def test_simple_river_tree():
    t = river_tree("50%", "60%", 100, 100)
    assert acts(t) == [(A.Check, 0), (A.Bet, 50), (A.AllIn, 100)]


def test_bet_clamped_to_allin():
    t = river_tree("50%", "60%", 100, 60)
    assert acts(t) == [(A.Check, 0), (A.AllIn, 60)]


def test_river_raise_forced_allin():
    t = river_tree("50%", "60%", 100, 200)
    assert acts(t, [(A.Bet, 50)]) == [(A.Fold, 0), (A.Call, 0), (A.AllIn, 200)]


def test_river_raise_survives_threshold():
    t = river_tree("33%", "60%", 100, 500)
    a = acts(t, [(A.Bet, 33)])
    assert (A.Raise, 133) in a and (A.Fold, 0) in a and (A.Call, 0) in a


def test_geometric():
    t = river_tree("e", "", 100, 100)
    assert acts(t) == [(A.Check, 0), (A.AllIn, 100)]


def test_merging():
    assert acts(river_tree("50%, 55%", "", 100, 300)) == [(A.Check, 0), (A.Bet, 55)]
    a = acts(river_tree("33%, 100%", "", 100, 300))
    assert (A.Bet, 33) in a and (A.Bet, 100) in a


def test_additive_and_prev_bet_relative():
    assert (A.Bet, 100) in acts(river_tree("100c", "", 100, 500))
    assert (A.Raise, 150) in acts(river_tree("50%", "3x", 100, 500), [(A.Bet, 50)])


def test_tiny_stack_allin_only():
    assert acts(river_tree("50%", "", 100, 5)) == [(A.Check, 0), (A.AllIn, 5)]


def test_terminal_lines_and_invalid_history():
    t = river_tree("50%", "60%", 100, 100)
    assert acts(t, [(A.Check, 0)]) == [(A.Check, 0), (A.Bet, 50), (A.AllIn, 100)]
    assert acts(t, [(A.Check, 0), (A.Check, 0)]) == []          # showdown
    assert acts(t, [(A.Bet, 50), (A.Fold, 0)]) == []            # fold
    assert acts(t, [(A.Bet, 50), (A.Call, 0)]) == []            # call on the river
    with pytest.raises(P.PsrError):
        acts(t, [(A.Bet, 51)])


def three_street(stack, donk=None, depth=None):
    s = [("50%", ""), ("50%", "")]
    return P.ActionTree(P.TreeConfig(initial_state=P.BoardState.Flop, starting_pot=100, effective_stack=stack,
                                     flop_bet_sizes=s, turn_bet_sizes=s, river_bet_sizes=s, turn_donk_sizes=donk,
                                     add_allin_threshold=1.5, force_allin_threshold=0.15, merging_threshold=0.12,
                                     depth_limit=depth))


def test_flop_tree_streets_and_counts():
    t = three_street(1000)
    # check-check on every street walks flop -> turn -> river -> showdown
    assert (A.Check, 0) in acts(t, [(A.Check, 0), (A.Check, 0)])
    h = [(A.Check, 0)] * 4
    assert (A.Check, 0) in acts(t, h)
    assert acts(t, [(A.Check, 0)] * 6) == []
    c = three_street(200).count_nodes()
    assert c[0] > 0 and c[1] > 0 and c[2] > 0
    assert river_tree("50%", "60%", 100, 200).count_nodes()[:2] == [0, 0]


def test_turn_donk_bet():
    t = three_street(1000, donk="75%")
    assert (A.Bet, 50) in acts(t, [(A.Check, 0)])
    a = acts(t, [(A.Check, 0), (A.Bet, 50), (A.Call, 0)])
    assert (A.Check, 0) in a and (A.Bet, 150) in a


def test_depth_limit_zero_makes_boundaries():
    s = [("50%", ""), ("50%", "")]
    cfg = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=100, effective_stack=200, turn_bet_sizes=s,
                       river_bet_sizes=s, add_allin_threshold=1.5, force_allin_threshold=0.15,
                       merging_threshold=0.12, depth_limit=0)
    t = P.ActionTree(cfg)
    assert acts(t, [(A.Check, 0), (A.Check, 0)]) == []
    assert acts(t, [(A.Bet, 50), (A.Call, 0)]) == []
    assert t.count_nodes()[2] == 0


def test_basic_example_actions():
    """examples/basic.rs:104-122: turn tree, bets "60%, e, a", raise "2.5x"."""
    s = [("60%, e, a", "2.5x"), ("60%, e, a", "2.5x")]
    cfg = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=200, effective_stack=900, flop_bet_sizes=s,
                       turn_bet_sizes=s, river_bet_sizes=s, river_donk_sizes="50%", add_allin_threshold=1.5,
                       force_allin_threshold=0.15, merging_threshold=0.1)
    t = P.ActionTree(cfg)
    assert acts(t) == [(A.Check, 0), (A.Bet, 120), (A.Bet, 216), (A.AllIn, 900)]
    assert acts(t, [(A.Bet, 120)]) == [(A.Fold, 0), (A.Call, 0), (A.Raise, 300)]


def test_max_raises_extension():
    # unlimited = reference behaviour; a cap of 1 removes re-raises only
    free = river_tree("50%", "3x", 100, 1000, add=0.0, force=0.0, merge=0.0)
    capped = river_tree("50%", "3x", 100, 1000, add=0.0, force=0.0, merge=0.0, max_raises=1)
    assert acts(free, [(A.Bet, 50)]) == acts(capped, [(A.Bet, 50)]) == [(A.Fold, 0), (A.Call, 0), (A.Raise, 150)]
    assert acts(free, [(A.Bet, 50), (A.Raise, 150)]) == [(A.Fold, 0), (A.Call, 0), (A.Raise, 450)]
    assert acts(capped, [(A.Bet, 50), (A.Raise, 150)]) == [(A.Fold, 0), (A.Call, 0)]


def test_bench_river_tree_shape():
    """The 1000-iteration river config: 21 nodes, 8 decisions, 20 edges."""
    t = river_tree("50%, 100%, a", "2.5x, a", 100, 100, add=0.0, force=0.0, merge=0.0, max_raises=1)
    assert acts(t) == [(A.Check, 0), (A.Bet, 50), (A.AllIn, 100)]
    assert acts(t, [(A.Bet, 50)]) == [(A.Fold, 0), (A.Call, 0), (A.AllIn, 100)]
    assert t.count_nodes() == [0, 0, 21]
