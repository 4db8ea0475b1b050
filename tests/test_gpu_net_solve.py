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
