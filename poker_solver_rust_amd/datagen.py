"""River training-data generation in the reference's distribution: the
situation sampler (cfvnet/src/datagen/sampler.rs:30-206) and DeepStack
R(S, p) ranges (cfvnet/src/datagen/range_gen.rs:19-170), solved by the batched
device solver and written as the reference's training records (records.py).

The random streams are numpy's, not the reference's ChaCha8, so samples are
drawn from the same distributions rather than equal to the reference's for a
seed.  Hand strengths on 4-card boards are averaged over the river cards
with this library's evaluator (a rank index in the sorted 7-card table); the
reference averages its own bit-packed encoding, so the orderings of turn
boards can differ slightly (5-card boards give identical ordinal ranks with
any correct evaluator).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .range_solver import hand_evaluate
from .workloads import Subgame, _pair_cards

NUM_COMBOS = 1326


@dataclass
class DatagenConfig:
    """The fields of cfvnet's DatagenConfig the sampler reads (config.rs:142-240),
    with its defaults; bet sizes are the one-depth default (config.rs:27-31):
    bets 25% / 50% / 100% / all-in, no raises."""
    pot_intervals: List[Tuple[int, int]] = field(default_factory=lambda: [(4, 20), (20, 80), (80, 200), (200, 400)])
    spr_intervals: Optional[List[Tuple[float, float]]] = None
    initial_stack: int = 200
    bet: str = "25%, 50%, 100%, a"
    raise_: str = ""
    solver_iterations: int = 1000


# ------------------------------------------------------------ strengths --


def _ordinal(raw: np.ndarray, valid: np.ndarray) -> np.ndarray:
    """range_gen.rs:93-112 raw_to_ordinal: ascending sort of the valid combos
    by raw strength; equal strengths share the rank of the first of them
    (position + 1); blocked combos 0."""
    idx = np.flatnonzero(valid)
    order = idx[np.lexsort((idx, raw[idx]))]
    vals = raw[order]
    ranks = np.zeros(NUM_COMBOS, np.uint16)
    start = np.r_[0, np.flatnonzero(vals[1:] != vals[:-1]) + 1]
    first = np.zeros(len(order), np.int64)
    first[start] = start
    first = np.maximum.accumulate(first)
    ranks[order] = (first + 1).astype(np.uint16)
    return ranks


def compute_hand_strengths(board: Sequence[int]) -> np.ndarray:
    """range_gen.rs:19-91: ordinal strength of every combo (higher is
    stronger, 0 for combos blocked by the board); on a 4-card board the raw
    strength is the mean over the unblocked river cards."""
    board = [int(c) for c in board]
    if len(board) not in (4, 5):
        raise ValueError("board must have 4 or 5 cards")
    c1, c2 = _pair_cards()
    bset = set(board)
    valid = np.array([(int(a) not in bset) and (int(b) not in bset) for a, b in zip(c1, c2)])
    raw = np.zeros(NUM_COMBOS, np.float64)
    for i in np.flatnonzero(valid):
        a, b = int(c1[i]), int(c2[i])
        if len(board) == 5:
            raw[i] = hand_evaluate(board + [a, b]) + 1
        else:
            rivers = [r for r in range(52) if r not in bset and r != a and r != b]
            raw[i] = sum(hand_evaluate(board + [r, a, b]) + 1 for r in rivers) / len(rivers)
    return _ordinal(raw, valid)


# --------------------------------------------------------------- ranges --


def rsp_range(strengths: np.ndarray, rng: np.random.Generator) -> np.ndarray:
    """range_gen.rs:124-170 R(S, p): the valid combos sorted by ascending
    strength are split in halves recursively; a uniform draw divides the
    mass p of a group, the larger part going to the stronger half.  Sums to
    1; blocked combos 0."""
    valid = np.flatnonzero(strengths > 0)
    hands = valid[np.argsort(strengths[valid], kind="stable")]
    out = np.zeros(NUM_COMBOS, np.float32)
    stack = [(0, len(hands), 1.0)]
    while stack:
        lo, hi, p = stack.pop()
        n = hi - lo
        if n == 0 or p <= 0.0:
            continue
        if n == 1:
            out[hands[lo]] += np.float32(p)
            continue
        draw = rng.random() * p
        rem = p - draw
        mid = lo + n // 2
        # stronger half first, as the reference recurses (the draws' order)
        stack.append((lo, mid, min(draw, rem)))
        stack.append((mid, hi, max(draw, rem)))
    return out


# ------------------------------------------------------------- sampler --


def sample_board(n: int, rng: np.random.Generator) -> List[int]:
    """sampler.rs:111-128: n distinct cards, drawn one at a time."""
    used, out = set(), []
    while len(out) < n:
        c = int(rng.integers(0, 52))
        if c not in used:
            used.add(c)
            out.append(c)
    return out


def sample_pot(intervals, rng: np.random.Generator) -> int:
    """sampler.rs:132-136: a uniform interval, then a uniform pot in [lo, hi)."""
    lo, hi = intervals[int(rng.integers(0, len(intervals)))]
    return int(rng.integers(lo, hi))


def sample_pot_stack_by_spr(pot_intervals, spr_intervals, initial_stack: int,
                            rng: np.random.Generator) -> Tuple[int, int]:
    """sampler.rs:138-206: an SPR bucket uniformly, a target SPR in it, a pot
    bucket uniformly among those feasible for that SPR, a pot in it; the
    stack is round(SPR * pot) clamped to [5, initial - pot / 2], kept when
    the resulting SPR lands in the chosen bucket."""
    stack_f = float(initial_stack)
    for _ in range(10000):
        s_lo, s_hi = spr_intervals[int(rng.integers(0, len(spr_intervals)))]
        spr = s_lo + (s_hi - s_lo) * rng.random()
        pot_min = np.ceil(5.0 / spr) if spr > 0.0 else np.ceil(2.0 * (stack_f - 5.0))
        pot_max = stack_f / (spr + 0.5)
        f_lo, f_hi = int(max(pot_min, 1.0)), int(pot_max)
        if f_lo > f_hi:
            continue
        buckets = [(max(a, f_lo), min(b, f_hi + 1)) for a, b in pot_intervals if max(a, f_lo) < min(b, f_hi + 1)]
        if not buckets:
            continue
        b_lo, b_hi = buckets[int(rng.integers(0, len(buckets)))]
        pot = int(rng.integers(b_lo, b_hi))
        max_stack = initial_stack - pot // 2
        if max_stack < 5:
            continue
        stack = int(min(max(int(np.round(spr * pot)), 5), max_stack))
        if s_lo <= stack / pot < s_hi:
            return pot, stack
    raise ValueError("pot and SPR intervals are not feasible for the initial stack")


def sample_situation(cfg: DatagenConfig, board_size: int, rng: np.random.Generator) -> Subgame:
    """sampler.rs:30-60 sample_situation, as a solvable subgame: board, pot,
    effective stack (the SPR sampler when SPR intervals are configured, else
    a pot from the pot intervals and a stack uniform in [5, initial - pot/2]),
    two R(S, p) ranges on the board's strengths."""
    board = sample_board(board_size, rng)
    if cfg.spr_intervals:
        pot, stack = sample_pot_stack_by_spr(cfg.pot_intervals, cfg.spr_intervals, cfg.initial_stack, rng)
    else:
        pot = sample_pot(cfg.pot_intervals, rng)
        max_stack = cfg.initial_stack - pot // 2
        stack = 5 if max_stack < 5 else int(rng.integers(5, max_stack + 1))
    strengths = compute_hand_strengths(board)
    ranges = (rsp_range(strengths, rng), rsp_range(strengths, rng))
    return Subgame(board, ranges, pot, stack, cfg.bet, cfg.raise_, 0, cfg.solver_iterations)


def generate_river_records(n: int, cfg: Optional[DatagenConfig] = None, seed: int = 0, chunk: int = 4096,
                           target_pot_fraction: float = 0.003):
    """End-to-end river datagen: n sampled situations solved in batches on the
    device (the reference solve() schedule, stopping at
    `target_pot_fraction` of the pot), two training records per situation
    (records.records_from_batch).  Returns (records, situations)."""
    from .range_solver import GpuBatchSolver
    from .records import records_from_batch

    cfg = cfg or DatagenConfig()
    rng = np.random.default_rng(seed)
    recs, sits = [], []
    for c0 in range(0, n, chunk):
        subs = [sample_situation(cfg, 5, rng) for _ in range(min(chunk, n - c0))]
        games = [s.build() for s in subs]
        b = GpuBatchSolver(0)
        b.prepare_batch(games, cfg.solver_iterations, [target_pot_fraction * s.pot for s in subs])
        b.solve(cfg.solver_iterations, 5)
        recs += records_from_batch(b, subs)
        sits += subs
    return recs, sits
