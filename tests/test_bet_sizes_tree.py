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


# --------------------------------------------------------------------------
# action_tree.rs:1238-2363: Action / BoardState / constants, navigation
# (play, undo, back_to_root, apply_history, is_terminal_node,
# is_chance_node, total_bet_amount), invalid terminals, depth limits.

def ref_river(bet, raise_, pot, stack):
    """action_tree.rs:1543-1557 river_config."""
    return river_tree(bet, raise_, pot, stack)


def ref_three_street(stack, depth=None):
    s = [("50%", ""), ("50%", "")]
    return P.ActionTree(P.TreeConfig(initial_state=P.BoardState.Flop, starting_pot=100, effective_stack=stack,
                                     flop_bet_sizes=s, turn_bet_sizes=s, river_bet_sizes=s, add_allin_threshold=1.5,
                                     force_allin_threshold=0.15, merging_threshold=0.12, depth_limit=depth))


def ref_turn(depth):
    s = [("50%", ""), ("50%", "")]
    return P.ActionTree(P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=100, effective_stack=200,
                                     turn_bet_sizes=s, river_bet_sizes=s, add_allin_threshold=1.5,
                                     force_allin_threshold=0.15, merging_threshold=0.12, depth_limit=depth))


def test_action_display_simple():
    assert P.action_to_string((A.None_, 0)) == "None"
    assert P.action_to_string((A.Fold, 0)) == "Fold"
    assert P.action_to_string((A.Check, 0)) == "Check"
    assert P.action_to_string((A.Call, 0)) == "Call"


def test_action_display_with_amount():
    assert P.action_to_string((A.Bet, 100)) == "Bet(100)"
    assert P.action_to_string((A.Raise, 250)) == "Raise(250)"
    assert P.action_to_string((A.AllIn, 1000)) == "AllIn(1000)"


def test_action_display_chance():
    assert P.action_to_string((A.Chance, 0)) == "Chance(2c)"
    assert P.action_to_string((A.Chance, 51)) == "Chance(As)"


def test_action_display_chance_invalid_card():
    assert P.action_to_string((A.Chance, 255)) == "Chance(255)"


def test_action_ordering():
    chain = [(A.None_, 0), (A.Fold, 0), (A.Check, 0), (A.Call, 0), (A.Bet, 1), (A.Bet, 2), (A.Bet, 100),
             (A.Raise, 1), (A.Raise, 100), (A.AllIn, 1), (A.AllIn, 100), (A.Chance, 0)]
    assert all(a < b for a, b in zip(chain, chain[1:]))


def test_action_eq_and_hash():
    assert len({(A.Bet, 100), (A.Bet, 100), (A.Raise, 100)}) == 2


def test_board_state_display_order_repr_default():
    assert [str(s) for s in (P.BoardState.Flop, P.BoardState.Turn, P.BoardState.River)] == ["Flop", "Turn", "River"]
    assert P.BoardState.Flop < P.BoardState.Turn < P.BoardState.River
    assert [int(s) for s in P.BoardState] == [0, 1, 2]
    assert P.TreeConfig().initial_state == P.BoardState.Flop


def test_player_constants():
    assert (P.PLAYER_OOP, P.PLAYER_IP, P.PLAYER_CHANCE, P.PLAYER_MASK) == (0, 1, 2, 3)
    assert (P.PLAYER_CHANCE_FLAG, P.PLAYER_TERMINAL_FLAG, P.PLAYER_FOLD_FLAG) == (4, 8, 24)
    assert P.PLAYER_FOLD_FLAG & P.PLAYER_TERMINAL_FLAG != 0


def test_depth_boundary_flag_constant():
    assert P.PLAYER_DEPTH_BOUNDARY_FLAG & P.PLAYER_TERMINAL_FLAG != 0
    assert P.PLAYER_DEPTH_BOUNDARY_FLAG & P.PLAYER_CHANCE_FLAG == 0
    assert P.PLAYER_DEPTH_BOUNDARY_FLAG != P.PLAYER_FOLD_FLAG


def test_tree_config_default():
    c = P.TreeConfig()
    assert (c.starting_pot, c.effective_stack, c.rake_rate, c.rake_cap) == (0, 0, 0.0, 0.0)
    assert (c.add_allin_threshold, c.force_allin_threshold, c.merging_threshold) == (0.0, 0.0, 0.0)
    assert c.turn_donk_sizes is None and c.river_donk_sizes is None and c.depth_limit is None


def test_tree_config_custom():
    s = [("60%, e, a", "2.5x"), ("60%, e, a", "2.5x")]
    c = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=200, effective_stack=900, rake_rate=0.05,
                     rake_cap=30.0, turn_bet_sizes=s, river_bet_sizes=s, river_donk_sizes="50%",
                     add_allin_threshold=1.5, force_allin_threshold=0.15, merging_threshold=0.1)
    t = P.ActionTree(c)
    assert t.config.starting_pot == 200 and t.config.effective_stack == 900
    assert abs(t.config.rake_rate - 0.05) < 1e-15 and abs(t.config.rake_cap - 30.0) < 1e-15
    assert t.config.river_donk_sizes is not None


def test_action_tree_default_accessors():
    t = ref_river("50%", "60%", 100, 100)
    assert t.history() == [] and t.num_invalid_terminals() == 0


def test_check_config_rejects_negative_pot():
    with pytest.raises(P.PsrError):
        P.ActionTree(P.TreeConfig(starting_pot=-10, effective_stack=100))


def test_river_check_check_is_terminal():
    t = ref_river("50%", "60%", 100, 100)
    t.play((A.Check, 0))
    assert (A.Check, 0) in t.available_actions()
    t.play((A.Check, 0))
    assert t.is_terminal_node()


def test_river_bet_fold_is_terminal():
    t = ref_river("50%", "60%", 100, 100)
    t.play((A.Bet, 50))
    a = t.available_actions()
    assert (A.Fold, 0) in a and (A.Call, 0) in a
    t.play((A.Fold, 0))
    assert t.is_terminal_node()


def test_river_bet_call_is_terminal():
    t = ref_river("50%", "60%", 100, 100)
    t.play((A.Bet, 50))
    t.play((A.Call, 0))
    assert t.is_terminal_node()


def test_play_and_undo():
    t = ref_river("50%", "60%", 100, 100)
    assert t.history() == []
    t.play((A.Check, 0))
    assert t.history() == [(A.Check, 0)]
    t.undo()
    assert t.history() == []


def test_play_invalid_action():
    with pytest.raises(P.PsrError):
        ref_river("50%", "60%", 100, 100).play((A.Fold, 0))


def test_undo_at_root_is_err():
    with pytest.raises(P.PsrError):
        ref_river("50%", "60%", 100, 100).undo()


def test_back_to_root():
    t = ref_river("50%", "60%", 100, 100)
    t.play((A.Check, 0))
    t.play((A.Check, 0))
    t.back_to_root()
    assert t.history() == []


def test_apply_history():
    t = ref_river("50%", "60%", 100, 100)
    t.apply_history([(A.Check, 0), (A.Check, 0)])
    assert t.history() == [(A.Check, 0), (A.Check, 0)]
    assert t.is_terminal_node()


def test_apply_history_invalid():
    with pytest.raises(P.PsrError):
        ref_river("50%", "60%", 100, 100).apply_history([(A.Fold, 0)])


def test_total_bet_amount_at_root():
    t = ref_river("50%", "60%", 100, 100)
    assert t.total_bet_amount() == [0, 0]
    t.play((A.Bet, 50))
    assert t.total_bet_amount() == [50, 0]
    t.play((A.Call, 0))
    assert t.total_bet_amount() == [50, 50]


def test_total_bet_amount_raise_and_streets():
    """Raises pay the difference to the new amount plus the call; amounts
    accumulate over streets (chance nodes are stepped over)."""
    t = river_tree("50%", "3x", 100, 1000, add=0.0, force=0.0, merge=0.0)
    t.apply_history([(A.Bet, 50), (A.Raise, 150)])
    assert t.total_bet_amount() == [50, 150]
    t.play((A.Call, 0))
    assert t.total_bet_amount() == [150, 150]
    f = ref_three_street(1000)
    f.apply_history([(A.Bet, 50), (A.Call, 0), (A.Check, 0), (A.Bet, 100), (A.Call, 0)])
    assert f.total_bet_amount() == [150, 150]


def test_flop_tree_has_chance_nodes():
    t = ref_three_street(200)
    t.play((A.Check, 0))
    t.play((A.Check, 0))
    assert t.is_chance_node()
    assert (A.Check, 0) in t.available_actions()


def test_flop_through_all_streets():
    t = ref_three_street(1000)
    t.play((A.Check, 0))
    t.play((A.Check, 0))
    assert t.is_chance_node() and not t.is_terminal_node()
    t.play((A.Check, 0))
    t.play((A.Check, 0))
    assert t.is_chance_node()
    t.play((A.Check, 0))
    t.play((A.Check, 0))
    assert t.is_terminal_node()


def test_river_raise_actions():
    t = ref_river("50%", "60%", 100, 200)
    t.play((A.Bet, 50))
    a = t.available_actions()
    assert (A.Fold, 0) in a and (A.Call, 0) in a and (A.AllIn, 200) in a


def test_prev_bet_relative_raise_navigation():
    t = ref_river("50%", "3x", 100, 500)
    t.play((A.Bet, 50))
    assert (A.Raise, 150) in t.available_actions()


def test_turn_donk_bet_navigation():
    s = [("50%", ""), ("50%", "")]
    t = P.ActionTree(P.TreeConfig(initial_state=P.BoardState.Flop, starting_pot=100, effective_stack=1000,
                                  flop_bet_sizes=s, turn_bet_sizes=s, turn_donk_sizes="75%", add_allin_threshold=1.5,
                                  force_allin_threshold=0.15, merging_threshold=0.12))
    t.play((A.Check, 0))
    assert (A.Bet, 50) in t.available_actions()
    t.back_to_root()
    t.play((A.Check, 0))
    t.play((A.Bet, 50))
    t.play((A.Call, 0))
    assert t.is_chance_node()
    a = t.available_actions()
    assert (A.Check, 0) in a and (A.Bet, 150) in a


def test_no_invalid_terminals():
    assert ref_river("50%", "60%", 100, 100).num_invalid_terminals() == 0
    assert ref_three_street(200).num_invalid_terminals() == 0


def test_count_action_nodes_river_exact():
    c = ref_river("50%", "60%", 100, 200).count_nodes()
    assert c[0] == 0 and c[1] == 0 and c[2] > 0


def test_depth_limit_none_does_not_affect_tree():
    t = ref_turn(None)
    t.play((A.Check, 0))
    t.play((A.Check, 0))
    assert t.is_chance_node()


def test_depth_limit_zero_blocks_all_transitions():
    t = ref_turn(0)
    assert not t.is_terminal_node() and not t.is_chance_node()
    t.play((A.Check, 0))
    t.play((A.Check, 0))
    assert t.is_terminal_node() and not t.is_chance_node()


def test_depth_limit_zero_bet_call_is_terminal():
    t = ref_turn(0)
    t.play((A.Bet, 50))
    t.play((A.Call, 0))
    assert t.is_terminal_node()


def test_depth_limit_one_allows_one_transition():
    t = ref_three_street(1000, depth=1)
    t.play((A.Check, 0))
    t.play((A.Check, 0))
    assert t.is_chance_node()
    t.play((A.Check, 0))
    t.play((A.Check, 0))
    assert t.is_terminal_node() and not t.is_chance_node()
