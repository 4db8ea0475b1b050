"""Situation sampler and R(S, p) ranges (cfvnet/src/datagen/sampler.rs,
range_gen.rs), mirroring the reference's own tests: ranges sum to one, are
non-negative and zero on board-blocked combos, stronger hands get more reach
on average, determinism per seed, pots inside the configured intervals,
stacks in [5, initial - pot / 2], the SPR sampler's buckets and bounds; hand
strengths against a brute-force ranking."""
import numpy as np

import poker_solver_rust_amd as P
from poker_solver_rust_amd import datagen as D
from poker_solver_rust_amd.workloads import _pair_cards

C1, C2 = _pair_cards()
BOARD = [0, 13, 26, 39, 7]


def blocked(board):
    return np.array([(int(a) in board) or (int(b) in board) for a, b in zip(C1, C2)])


def test_ordinal_ranks_share_first_position():
    raw = np.zeros(1326)
    valid = np.zeros(1326, bool)
    valid[[3, 5, 9, 11]] = True
    raw[[3, 5, 9, 11]] = [7.0, 2.0, 7.0, 4.0]
    r = D._ordinal(raw, valid)
    assert r[5] == 1 and r[11] == 2 and r[3] == 3 and r[9] == 3 and r[0] == 0


def test_strengths_5_card_brute_force():
    s = D.compute_hand_strengths(BOARD)
    bm = blocked(BOARD)
    assert np.all(s[bm] == 0) and np.all(s[~bm] > 0)
    raw = {i: P.hand_evaluate(BOARD + [int(C1[i]), int(C2[i])]) for i in np.flatnonzero(~bm)}
    for i in list(raw)[:40]:
        # rank = 1 + number of valid combos strictly weaker
        assert s[i] == 1 + sum(1 for v in raw.values() if v < raw[i])


def test_strengths_4_card_board():
    board = BOARD[:4]
    s = D.compute_hand_strengths(board)
    bm = blocked(board)
    assert np.all(s[bm] == 0) and np.all(s[~bm] > 0)


def test_rsp_ranges_valid_and_strength_correlated():
    rng = np.random.default_rng(42)
    s = D.compute_hand_strengths(BOARD)
    bm = blocked(BOARD)
    valid = np.flatnonzero(~bm)
    order = valid[np.argsort(s[valid], kind="stable")]
    weak, strong = order[:len(order) // 2], order[len(order) // 2:]
    acc_w = acc_s = 0.0
    for _ in range(200):
        r = D.rsp_range(s, rng)
        assert abs(float(r.sum()) - 1.0) < 1e-4
        assert np.all(r >= 0) and np.all(r[bm] == 0)
        acc_w += r[weak].sum()
        acc_s += r[strong].sum()
    assert acc_s > acc_w   # range_gen.rs strong_hands_have_higher_mean_reach


def test_sampler_bounds_and_determinism():
    cfg = D.DatagenConfig(pot_intervals=[(4, 20), (20, 80), (80, 200), (200, 400)], solver_iterations=100)
    rng = np.random.default_rng(7)
    for _ in range(30):
        s = D.sample_situation(cfg, 5, rng)
        assert len(set(s.board)) == 5 and all(0 <= c < 52 for c in s.board)
        assert any(lo <= s.pot < hi for lo, hi in cfg.pot_intervals)
        assert 5 <= s.stack <= max(5, cfg.initial_stack - s.pot // 2)
        for p in range(2):
            assert abs(float(s.ranges[p].sum()) - 1.0) < 1e-4 and np.all(s.ranges[p][blocked(s.board)] == 0)
    a = D.sample_situation(cfg, 5, np.random.default_rng(99))
    b = D.sample_situation(cfg, 5, np.random.default_rng(99))
    assert a.board == b.board and a.pot == b.pot and a.stack == b.stack
    assert np.array_equal(a.ranges[0], b.ranges[0]) and np.array_equal(a.ranges[1], b.ranges[1])


def test_spr_sampler():
    """sampler.rs spr_stratified_{covers_all_buckets, respects_bounds,
    actual_spr_in_bucket}."""
    pots = [(4, 20), (20, 80), (80, 200)]
    sprs = [(0.0, 0.5), (0.5, 1.5), (1.5, 4.0), (4.0, 8.0), (8.0, 50.0)]
    rng = np.random.default_rng(42)
    hits = np.zeros(len(sprs), int)
    for _ in range(1000):
        pot, stack = D.sample_pot_stack_by_spr(pots, sprs, 200, rng)
        assert 5 <= stack <= 200 - pot // 2
        assert any(lo <= pot < hi for lo, hi in pots)
        spr = stack / pot
        k = [i for i, (lo, hi) in enumerate(sprs) if lo <= spr < hi]
        assert len(k) == 1
        hits[k[0]] += 1
    assert np.all(hits > 0)
