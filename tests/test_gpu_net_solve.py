"""The value net in its place: river subgames solved by the batched device
solver are the training data, the net learns the exact-minus-check-down
correction, and on held-out subgames the net-backed leaf value is closer to
the exact river solve than the check-down value alone (the method without
the net).  Then a short net-backed depth-limited turn solve runs end to end.

Values are pot-relative (v_i = E[payoff of i] / (pot / 2)); the error is the
mean absolute difference over the combos present in each subgame.
"""
import numpy as np
import pytest

from poker_solver_rust_amd import net_solve as N
from poker_solver_rust_amd import range_solver as R
from poker_solver_rust_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def turn():
    return W.turn_uniform01(20260)


@pytest.fixture(scope="module")
def trained(turn):
    game = R.PostFlopGame.with_config(turn.card_config(), R.ActionTree(N.depth_limited_tree(turn)))
    pots = sorted({int(game.boundary_pot(b)) for b in range(game.num_boundary_nodes())})
    checkdown = N.CheckdownLeaves(turn.board[:4])
    ts = N.river_training_set(turn, pots, 6144, seed=3, iterations=200, checkdown=checkdown, log=lambda s: None)
    net, losses, mae = N.train_value_net(ts, epochs=30, batch=256, seed=1, holdout=0.1)
    return ts, net, losses, mae, checkdown


def test_net_beats_checkdown_on_held_out_river_solves(trained):
    ts, net, losses, mae, _ = trained
    nh = max(1, int(len(ts.x) * 0.1))
    m = ts.mask[:nh]
    # the targets are exact - check-down, so predicting 0 is the check-down value
    err_checkdown = float((np.abs(ts.y[:nh]) * m).sum() / m.sum())
    assert losses[-1] < losses[0]
    assert np.isfinite(mae) and mae < 0.8 * err_checkdown, (mae, err_checkdown)


def test_net_backed_depth_limited_solve_runs(turn, trained):
    _, net, _, _, checkdown = trained
    leaves = N.NetLeaves(net, checkdown)
    game, b, t_leaf, t_dcfr = N.solve_depth_limited(turn, leaves, iterations=20, refresh=10, log=lambda s: None)
    cfv = b.root_cfvalues()[0]
    assert all(np.all(np.isfinite(v)) for v in cfv)
    # leaves were set at every boundary of both players
    for k in range(game.num_boundary_nodes()):
        for p in range(2):
            assert np.all(np.isfinite(b.boundary_reach(0, k, p)))
    # the batched leaf path equals the per-boundary call
    idx = [N.hand_index(game, p) for p in range(2)]
    items = []
    for k in range(min(3, game.num_boundary_nodes())):
        reach = []
        for p in range(2):
            full = np.zeros(1326)
            full[idx[p]] = b.boundary_reach(0, k, p)
            reach.append(full)
        items.append((turn.board[:4], reach, float(game.boundary_pot(k)), N.boundary_stacks(turn, game)[k]))
    batched = leaves.values_for(items)
    for it, (v0, v1) in zip(items, batched):
        s0, s1 = leaves(*it)
        assert np.allclose(v0, s0, atol=1e-4) and np.allclose(v1, s1, atol=1e-4)


@pytest.fixture(scope="module")
def c5_ensemble(turn):
    """Config C5's pipeline at test size: a check-down depth-limited solve
    records the boundary reaches its iterations visit; 16384 river subgames
    (half on those reaches with noise, half boundary-like) solved by the
    batched device solver (solve() schedule, cap 300); three value nets
    trained on the device on the exact-minus-check-down targets (seeds 20260,
    7, 11)."""
    game = R.PostFlopGame.with_config(turn.card_config(), R.ActionTree(N.depth_limited_tree(turn)))
    pots = sorted({int(game.boundary_pot(b)) for b in range(game.num_boundary_nodes())})
    chk = N.CheckdownLeaves(turn.board[:4])
    snaps = []
    _, b_chk, _, _ = N.solve_depth_limited(turn, chk, 300, 10, log=lambda s: None, snapshots=snaps)
    sits = [it for items in snaps for it in items]
    ts = N.river_training_set(turn, pots, 16384, seed=3, iterations=300, log=lambda s: None, situations=sits,
                              checkdown=chk)
    nets = [N.train_value_net(ts, epochs=30, seed=s)[0] for s in (20260, 7, 11)]
    dev = N.DeviceCheckdownLeaves(turn.board[:4], game.num_boundary_nodes())
    return dev, b_chk, nets


def test_net_backed_turn_solve_within_one_percent_of_the_pot(turn, c5_ensemble):
    """Config C5's target on the seeded turn subgame: the depth-limited turn
    solve with the value nets at the river deal (300 iterations, leaves
    refreshed every 10) lands within 1% of the pot of the exact full turn
    solve in game value, and closer than the same solve with check-down
    leaves (the method without the nets).  Game value: each player's
    range-weighted root expected value (query.rs:639-666).  Measured on the
    device: ensemble 0.46%, check-down 1.48% (the spread over training seeds
    and data sets is recorded with the method)."""
    dev, b_chk, nets = c5_ensemble
    _, b_net, _, _ = N.solve_depth_limited(turn, N.NetLeaves(nets, dev), 300, 10, log=lambda s: None)
    g_ex = turn.build()
    b_ex = R.GpuBatchSolver(0)
    b_ex.prepare_batch([g_ex], 300)
    b_ex.run_iterations(0, 300)

    def game_value_error(b):
        e = N.root_ev_error(b_ex, b, g_ex, turn.pot)
        return max(e["p0_value_diff_pot"], e["p1_value_diff_pot"])

    err_net, err_chk = game_value_error(b_net), game_value_error(b_chk)
    print(f"C5: net-backed game value error {err_net * 100:.3f}% of the pot, check-down {err_chk * 100:.3f}%")
    assert err_net < 0.01
    assert err_net < err_chk


def test_device_checkdown_equals_host_checkdown(turn):
    """The check-down leaf values from the device showdown evaluator
    (DeviceCheckdownLeaves) against the vectorised host ones on the boundary
    reaches of a short check-down depth-limited solve: per river and mixed
    over the rivers."""
    chk = N.CheckdownLeaves(turn.board[:4])
    game, b, _, _ = N.solve_depth_limited(turn, chk, 20, 10, log=lambda s: None)
    items = N.boundary_items(turn, game, b)
    dev = N.DeviceCheckdownLeaves(turn.board[:4], len(items))
    r0 = np.stack([it[1][0] for it in items])
    r1 = np.stack([it[1][1] for it in items])
    host = chk.per_river_rows(r0, r1)
    got = dev.per_river_rows(r0, r1)
    scale = float(np.abs(host).max())
    assert scale > 0 and float(np.abs(got - host).max()) <= 1e-5 * scale
    for (h0, h1), (d0, d1) in zip(chk.values_for(items), dev.values_for(items)):
        assert np.allclose(d0, h0, atol=1e-5 * scale) and np.allclose(d1, h1, atol=1e-5 * scale)
