"""Host-side checks of the device layout (no GPU): tiles, levels, child
placement, state and scan-table bounds for river, turn, flop and batch games."""
import pytest

import poker_solver_rust_amd as P
from poker_solver_rust_amd import workloads as W


def test_river_is_one_tile():
    info = W.river_uniform(0).build().validate_layout()
    assert info["tiles"] == 1 and info["nodes"] == 21 and info["state"] == 21620


def test_turn_tiles_cover_48_rivers_per_chance_node():
    g = W.turn_uniform01(0).build()
    info = g.validate_layout()
    assert (info["tiles"] - 1) % 48 == 0 and info["tiles"] > 48
    assert info["nodes"] == g.num_nodes() and info["state"] == g.num_storage()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_river_batch_layouts(seed):
    for s in W.river_batch(40, seed=seed):
        g = s.build()
        info = g.validate_layout()
        assert info["tiles"] == 1 and info["state"] == g.num_storage()


def test_turn_with_allin_chance_children():
    """All-in on the turn deals rivers straight into showdowns: subtree tiles
    that are a single terminal."""
    cc = P.CardConfig(range=(P.Range.parse("AA,KK"), P.Range.parse("QQ,JJ")), flop=P.flop_from_str("2c7d9h"),
                      turn=P.card_from_str("Ts"))
    s = [("a", ""), ("a", "")]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=100, effective_stack=100, turn_bet_sizes=s,
                      river_bet_sizes=s)
    g = P.PostFlopGame.with_config(cc, P.ActionTree(tc))
    info = g.validate_layout()
    assert info["tiles"] > 48


def test_flop_games_lay_out_in_three_tile_levels():
    """Flop-root games: the root tile (flop betting), one tile per turn card
    below each of the 3 flop chance nodes (the turn betting down to its 3
    river deals), one tile per river card below each of those; on a rainbow
    flop no two cards are suit-isomorphic: 1 + 3 * 49 + 3 * 49 * 3 * 48."""
    s = [("50%", ""), ("50%", "")]
    cc = P.CardConfig(range=(P.Range.parse("AA"), P.Range.parse("KK")), flop=P.flop_from_str("2c7d9h"))
    tc = P.TreeConfig(initial_state=P.BoardState.Flop, starting_pot=100, effective_stack=100, flop_bet_sizes=s,
                      turn_bet_sizes=s, river_bet_sizes=s)
    g = P.PostFlopGame.with_config(cc, P.ActionTree(tc))
    info = g.validate_layout()
    assert info["tiles"] == 1 + 3 * 49 + 3 * 49 * 3 * 48
    assert info["nodes"] == g.num_nodes() and info["state"] == g.num_storage()
