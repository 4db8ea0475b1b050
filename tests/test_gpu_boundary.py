"""Depth-boundary leaves on the GPU, through the C ABI, against the oracle.

Turn trees cut at the river deal (depth_limit 0): seeded leaf values, scaled
(evaluation.rs:544-604) on even ordinals and raw (evaluation.rs:606-630) on
odd ones; regrets, strategy sums, the recorded boundary reach
(evaluation.rs:91-99) and exploitability compared with the oracle after every
iteration, and the batch-wide update_leaf_cfvs (batch.rs:557-585) against
per-boundary calls.
"""
import numpy as np
import pytest

import oracle_binding as O
import poker_solver_rust_amd as P
from poker_solver_rust_amd import workloads as W

pytestmark = pytest.mark.gpu


def boundary_game(oop="AA,KK,QQ,AKs", ip="QQ-JJ,AQs,AJs", bets=("50%, a", "2x"), turn="8d"):
    cc = P.CardConfig(range=(P.Range.parse(oop), P.Range.parse(ip)), flop=P.flop_from_str("Qs Jh 2c"),
                      turn=P.card_from_str(turn))
    s = [bets, bets]
    tc = P.TreeConfig(initial_state=P.BoardState.Turn, starting_pot=100, effective_stack=200, turn_bet_sizes=s,
                      river_bet_sizes=s, depth_limit=0)
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def values(g, seed):
    rng = np.random.default_rng(seed)
    return [[rng.uniform(-1, 1, g.num_private_hands(q)).astype(np.float32) for q in range(2)]
            for _ in range(g.num_boundary_nodes())]


@pytest.mark.parametrize("make", [boundary_game, lambda: boundary_game(bets=("33%, 75%, a", "2.5x, a"))])
def test_boundary_stepwise_matches_oracle(make):
    g = make()
    nb = g.num_boundary_nodes()
    assert nb > 2
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 12)
    assert b.num_boundaries(0) == nb
    o = O.OracleGame(g.serialize())
    vals = values(g, 7)
    for k in range(nb):
        for q in range(2):
            b.set_boundary_cfvs(0, k, q, vals[k][q], raw=k % 2 == 1)
            o.set_boundary_cfvs(k, q, vals[k][q], raw=k % 2 == 1)
    for t in range(10):
        b.run_iterations(t, t + 1)
        o.solve_step(t)
        r, s = o.state()
        assert np.array_equal(b.regrets(0), r.astype(np.float32)), t
        assert np.array_equal(b.strategy_sum(0), s.astype(np.float32)), t
        for k in range(nb):
            for q in range(2):
                assert np.array_equal(b.boundary_reach(0, k, q), o.boundary_reach(k, q)), (t, k, q)
    assert abs(float(b.exploitability()[0]) - o.exploitability()) <= 1e-6 * max(1.0, abs(o.exploitability()))
    cf = b.root_cfvalues()[0]
    assert np.array_equal(cf[0], o.root_cfvalues(0).astype(np.float32))
    assert np.array_equal(cf[1], o.root_cfvalues(1).astype(np.float32))


def test_update_leaf_cfvs_equals_per_boundary_raw():
    """batch.rs:557-585 layout: per game, per ordinal, H_player values."""
    games = [boundary_game(), boundary_game(oop="AA,AKs", ip="KK,QQ", turn="3h")]
    a = P.GpuBatchSolver(0)
    a.prepare_batch(games, 6)
    c = P.GpuBatchSolver(0)
    c.prepare_batch(games, 6)
    p0, p1 = [], []
    for gi, g in enumerate(games):
        for k, (v0, v1) in enumerate(values(g, 11 + gi)):
            p0.append(v0)
            p1.append(v1)
            c.set_boundary_cfvs(gi, k, 0, v0, raw=True)
            c.set_boundary_cfvs(gi, k, 1, v1, raw=True)
    a.update_leaf_cfvs(np.concatenate(p0), np.concatenate(p1))
    a.run_iterations(0, 6)
    c.run_iterations(0, 6)
    for gi in range(2):
        assert np.array_equal(a.regrets(gi), c.regrets(gi))
        assert np.array_equal(a.strategy_sum(gi), c.strategy_sum(gi))
    with pytest.raises(P.PsrError):
        a.update_leaf_cfvs(np.concatenate(p0)[:-1], np.concatenate(p1))


def test_boundary_errors():
    g = boundary_game()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 4)
    with pytest.raises(P.PsrError):
        b.set_boundary_cfvs(0, g.num_boundary_nodes(), 0, np.zeros(g.num_private_hands(0), np.float32))
    with pytest.raises(P.PsrError):
        b.set_boundary_cfvs(0, 0, 0, np.zeros(g.num_private_hands(0) + 1, np.float32))
    with pytest.raises(P.PsrError):
        b.boundary_reach(1, 0, 0)


def _combo_rows(g, q, values_by_hand, stride=1326):
    """Values in game q's hand order placed at their card_pair_to_index slots."""
    out = np.zeros(stride, np.float32)
    for h, (c1, c2) in enumerate(g.private_cards(q)):
        out[P.card_pair_to_index(c1, c2)] = values_by_hand[h]
    return out


@pytest.mark.parametrize("stride", [1326, 0])
def test_leaf_injection_layout_matches_per_boundary_values(stride):
    """set_leaf_injection(node_ids, depths) + update_leaf_cfvs in the
    reference's [game][leaf][hand] layout (batch.rs:279-300, 557-585; leaves
    in a shuffled order, hands in the 1326-combo space or the game's own
    order) solves exactly like the same raw values given per boundary."""
    games = [boundary_game(), boundary_game(oop="AA,KK,AKs,AQs", ip="QQ-TT,AJs")]
    nb = games[0].num_boundary_nodes()
    assert games[1].num_boundary_nodes() == nb
    nodes = np.array(games[0].boundary_node_indices(), np.int32)
    assert nodes.tolist() == games[1].boundary_node_indices()   # the same tree
    depth = games[0].topology()[1]
    order = np.random.default_rng(3).permutation(nb)
    st = stride or max(g.num_private_hands(q) for g in games for q in range(2)) + 5
    vals = [values(g, 11 + i) for i, g in enumerate(games)]
    a = P.GpuBatchSolver(0)
    a.set_leaf_injection(nodes[order], depth[nodes[order]], st)
    a.prepare_batch(games, 10)
    flat = [np.zeros((len(games), nb, st), np.float32) for _ in range(2)]
    for gi, g in enumerate(games):
        for li, k in enumerate(order):
            for q in range(2):
                row = _combo_rows(g, q, vals[gi][k][q]) if stride == 1326 else np.pad(
                    vals[gi][k][q], (0, st - g.num_private_hands(q)))
                flat[q][gi, li] = row
    a.update_leaf_cfvs(flat[0].ravel(), flat[1].ravel())
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 10)
    for gi in range(len(games)):
        for k in range(nb):
            for q in range(2):
                b.set_boundary_cfvs(gi, k, q, vals[gi][k][q], raw=True)
    a.run_iterations(0, 10)
    b.run_iterations(0, 10)
    for gi in range(len(games)):
        assert np.array_equal(a.regrets(gi), b.regrets(gi))
        assert np.array_equal(a.strategy_sum(gi), b.strategy_sum(gi))


def test_leaf_injection_rejects_wrong_nodes_and_depths():
    g = boundary_game()
    nodes = np.array(g.boundary_node_indices(), np.int32)
    depth = g.topology()[1]
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 5)
    with pytest.raises(P.PsrError, match="depth"):
        b.set_leaf_injection(nodes, depth[nodes] + 1, 1326)
    with pytest.raises(P.PsrError, match="not a depth-boundary leaf"):
        b.set_leaf_injection([0], [0], 1326)
    with pytest.raises(P.PsrError, match="hand stride"):
        b.set_leaf_injection(nodes, depth[nodes], 3)


@pytest.mark.parametrize("make", [boundary_game, lambda: W.river_uniform(5).build()])
def test_download_reach_is_the_last_traversals_reach(make):
    """download_reach (batch.rs:537-546) after run_iterations: player 0's
    reach of player 1's last traversal at every node.  The root holds the
    initial weights; a chance node's children hold its reach times 1/44; a
    player-1 node passes its reach through; at a player-0 node the children's
    reaches are sigma * reach, so they sum to the node's reach; the
    1326-combo layout holds the same values."""
    g = make()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 6)
    b.run_iterations(0, 6)
    H0 = g.num_private_hands(0)
    r = b.download_reach(H0 + 3)
    assert np.all(r[:, H0:] == 0.0)
    r = r[:, :H0]
    par, _ = g.topology()
    nodes = list(range(g.num_nodes()))
    kids = {}
    for c in nodes[1:]:
        kids.setdefault(int(par[c]), []).append(c)
    assert np.array_equal(r[0], g.initial_weights(0))
    blob_nodes = O.OracleGame(g.serialize()).node_kinds()
    for n, cs in kids.items():
        kind, player = blob_nodes[n]
        if kind == "chance":
            for c in cs:
                assert np.array_equal(r[c], (r[n] * np.float32(1.0 / 44.0)).astype(np.float32))
        elif player == 1 or len(cs) == 1:
            for c in cs:
                assert np.array_equal(r[c], r[n])
        else:
            tot = np.sum([r[c] for c in cs], axis=0)
            assert np.allclose(tot, r[n], rtol=2e-6, atol=1e-30)
    rc = b.download_reach(1326)
    for n in (0, len(nodes) - 1):
        assert np.array_equal(rc[n], _combo_rows(g, 0, r[n]))


@pytest.mark.parametrize("make", [boundary_game, lambda: W.river_uniform(6).build()])
def test_strategy_sum_edges_layout(make):
    """extract_results' [edge][hand] layout (batch.rs:587-609): edge e is the
    action leading to node e + 1; a decision node's actions carry its A x H
    block of the storage-order strategy sums, chance edges are zero."""
    g = make()
    b = P.GpuBatchSolver(0)
    b.prepare_batch([g], 8)
    b.run_iterations(0, 8)
    S = b.strategy_sum(0)
    st = 1200
    E = b.extract_results(st)[0]
    par, _ = g.topology()
    assert E.shape == (g.num_nodes() - 1, st)
    blob_nodes = O.OracleGame(g.serialize()).node_kinds()
    off, seen = 0, set()
    for n in range(g.num_nodes()):
        kind, p = blob_nodes[n]
        if kind != "decision":
            continue
        cs = [c for c in range(1, g.num_nodes()) if par[c] == n]
        H = g.num_private_hands(p)
        for a, c in enumerate(cs):
            assert np.array_equal(E[c - 1, :H], S[off + a * H: off + (a + 1) * H])
            assert np.all(E[c - 1, H:] == 0.0)
            seen.add(c - 1)
        off += len(cs) * H
    assert off == len(S)
    for e in set(range(g.num_nodes() - 1)) - seen:
        assert np.all(E[e] == 0.0)
    Ec = b.extract_results(1326)[0]
    for e in list(seen)[:5]:
        kind, p = blob_nodes[int(par[e + 1])]
        assert np.array_equal(Ec[e], _combo_rows(g, p, E[e, :g.num_private_hands(p)]))
