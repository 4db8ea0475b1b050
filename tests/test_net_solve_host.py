"""Host-side pieces of the net-backed turn solve (encode, unblocked opponent
mass, the river mix, check-down leaf values), against brute force."""
import numpy as np
import pytest

import poker_solver_rust_amd as P
from poker_solver_rust_amd import net_solve as N
from poker_solver_rust_amd.workloads import _pair_cards, blocked

C1, C2 = _pair_cards()


def rand_reach(rng, board):
    r = rng.random(1326)
    r[rng.random(1326) < 0.3] = 0.0
    return np.where(blocked(board), 0.0, r)


def test_unblocked_mass_brute_force():
    rng = np.random.default_rng(0)
    board = [3, 17, 30, 44]
    r = rand_reach(rng, board)
    got = N.unblocked_mass(r)
    for i in rng.choice(1326, 60, replace=False):
        ref = sum(r[j] for j in range(1326) if not ({C1[i], C2[i]} & {C1[j], C2[j]}))
        assert abs(got[i] - ref) < 1e-9
    got_c = N.unblocked_mass(r, 7)
    for i in rng.choice(1326, 30, replace=False):
        ref = sum(r[j] for j in range(1326) if not ({C1[i], C2[i], 7} & {C1[j], C2[j]}))
        assert abs(got_c[i] - ref) < 1e-9


def test_river_mix_weights_sum_to_chance_factor():
    """sum_c R_i(c) = 44 R_i for every hand not on the board: each opponent
    combo compatible with i leaves 52 - 4 - 2 - 2 = 44 river cards."""
    rng = np.random.default_rng(1)
    board = [3, 17, 30, 44]
    r = rand_reach(rng, board)
    rivers = [c for c in range(52) if c not in board]
    den = np.zeros(1326)
    for c in rivers:
        den += np.where((C1 == c) | (C2 == c), 0.0, N.unblocked_mass(r, c))
    ok = ~blocked(board)
    assert np.allclose(den[ok], 44 * N.unblocked_mass(r)[ok])
    # a river-independent value mixes to itself
    v = rng.normal(0, 1, 1326)
    mixed = N.river_mix(np.repeat(v[None], len(rivers), 0), r, rivers)
    has = ok & (N.unblocked_mass(r) > 0)
    assert np.allclose(mixed[has], v[has])


def test_encode_layout():
    rng = np.random.default_rng(2)
    board = [3, 17, 30, 44, 9]
    r0, r1 = rand_reach(rng, board), rand_reach(rng, board)
    x = N.encode(r0, r1, board, 100.0, 300.0)[0]
    assert x.shape == (2706,)
    for k, r in enumerate((r0, r1)):
        seg = x[k * 1326:(k + 1) * 1326]
        assert np.allclose(seg[r > 0].mean(), 1.0) and np.all(seg[r == 0] == 0)
        assert np.allclose(seg / seg.max(), r / r.max(), atol=1e-6)
    assert sorted(np.flatnonzero(x[2652:2704])) == sorted(board)
    assert np.allclose(x[-2:], [0.25, 0.75])


def test_checkdown_values_brute_force():
    rng = np.random.default_rng(3)
    board4 = [3, 17, 30, 44]
    leaves = N.CheckdownLeaves(board4)
    k, c = 5, leaves.rivers[5]
    board = board4 + [c]
    r = rand_reach(rng, board)
    got = leaves.values(k, r)[0]
    strength = leaves.tables[k][0] * 0
    for i in range(1326):
        if leaves.tables[k][0][i]:
            strength[i] = P.hand_evaluate(board + [int(C1[i]), int(C2[i])]) + 1
    for i in rng.choice(np.flatnonzero(~blocked(board)), 25, replace=False):
        num = den = 0.0
        for j in np.flatnonzero(r > 0):
            if {C1[i], C2[i]} & {C1[j], C2[j]}:
                continue
            den += r[j]
            num += r[j] * np.sign(strength[i] - strength[j])
        assert abs(got[i] - (num / den if den > 0 else 0.0)) < 1e-9


def test_checkdown_batched_equals_per_boundary():
    """values_for over several boundaries at once equals one boundary at a
    time (the vectorised path the depth-limited solve uses)."""
    rng = np.random.default_rng(5)
    board4 = [3, 17, 30, 44]
    leaves = N.CheckdownLeaves(board4)
    items = [(board4, [rand_reach(rng, board4), rand_reach(rng, board4)], 100.0 + 50 * k, 200.0) for k in range(3)]
    batched = leaves.values_for(items)
    for it, (v0, v1) in zip(items, batched):
        rivers = leaves.rivers
        per = np.stack([leaves.per_river_rows(it[1][0][None], it[1][1][None])[p, :, 0] for p in range(2)])
        assert np.allclose(v0, N.river_mix(per[0], it[1][1], rivers))
        assert np.allclose(v1, N.river_mix(per[1], it[1][0], rivers))


def test_zero_sum_correction():
    """After the correction the weighted values of the two players cancel;
    values that already cancel (check-down equities) are left unchanged."""
    rng = np.random.default_rng(6)
    board = [3, 17, 30, 44, 9]
    r0, r1 = rand_reach(rng, board), rand_reach(rng, board)
    v0, v1 = rng.normal(0, 1, 1326), rng.normal(0, 1, 1326)
    a0, a1 = N.zero_sum(v0, v1, r0, r1)
    R0, R1 = N.unblocked_mass(r1), N.unblocked_mass(r0)
    tot = (r0 * R0 * a0).sum() + (r1 * R1 * a1).sum()
    assert abs(tot) < 1e-9 * (r0 * R0).sum()
    assert np.allclose(a0 - v0, (a0 - v0)[0]) and np.allclose(a1 - v1, (a1 - v1)[0])
    leaves = N.CheckdownLeaves(board[:4])
    k = leaves.rivers.index(9)
    c0, c1 = leaves.values(k, r1)[0], leaves.values(k, r0)[0]
    b0, b1 = N.zero_sum(c0, c1, r0, r1)
    assert np.allclose(b0, c0, atol=1e-9) and np.allclose(b1, c1, atol=1e-9)
