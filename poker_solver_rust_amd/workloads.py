"""Seeded synthetic workloads of the range-solver benchmark configurations.

river_uniform   river subgame, 5 seeded board cards, both ranges weight 1 on
                all 1081 unblocked combos, pot 100, stack 100, bets 50%/100%/
                all-in, one raise (2.5x or all-in), 1000 DCFR iterations
turn_uniform01  turn subgame, 4 seeded board cards, i.i.d. U[0,1] combo
                weights, pot 60, stack 200, bets 33%/75%/all-in, up to two
                raises at 3x, 500 iterations
river_batch     N independent river subgames: random board, Dirichlet(0.5)
                ranges, pot log-uniform [20, 400], stack-to-pot U[0.5, 10],
                1-3 sizes from {25,33,50,75,100,150}% plus all-in, raise cap
                U{1,2,3}; 300 iterations each
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from .range_solver import (NOT_DEALT, ActionTree, BoardState, CardConfig, PostFlopGame, Range, TreeConfig,
                           card_pair_to_index)

BATCH_SIZES = (25, 33, 50, 75, 100, 150)


def _board(rng: np.random.Generator, n: int) -> List[int]:
    return [int(c) for c in rng.choice(52, size=n, replace=False)]


def _blocked_mask(board: Sequence[int]) -> np.ndarray:
    m = np.zeros(1326, bool)
    for a in range(52):
        for b in range(a + 1, 52):
            if a in board or b in board:
                m[card_pair_to_index(a, b)] = True
    return m


_PAIRS = None


def _pair_cards():
    global _PAIRS
    if _PAIRS is None:
        c1 = np.zeros(1326, np.int64)
        c2 = np.zeros(1326, np.int64)
        k = 0
        # card_pair_to_index order: (0,1),(0,2),...,(0,51),(1,2),...
        for a in range(52):
            for b in range(a + 1, 52):
                c1[k], c2[k] = a, b
                k += 1
        _PAIRS = (c1, c2)
    return _PAIRS


def blocked(board: Sequence[int]) -> np.ndarray:
    c1, c2 = _pair_cards()
    bm = np.zeros(52, bool)
    bm[list(board)] = True
    return bm[c1] | bm[c2]


def _sizes(pcts: Sequence[int]) -> str:
    return ", ".join([f"{p}%" for p in pcts] + ["a"])


@dataclass
class Subgame:
    board: List[int]
    ranges: tuple
    pot: int
    stack: int
    bet: str
    raise_: str
    max_raises: int
    iterations: int

    def card_config(self) -> CardConfig:
        b = self.board
        return CardConfig(range=(Range(self.ranges[0]), Range(self.ranges[1])), flop=sorted(b[:3]), turn=b[3],
                          river=b[4] if len(b) > 4 else NOT_DEALT)

    def tree_config(self) -> TreeConfig:
        state = BoardState.River if len(self.board) == 5 else BoardState.Turn
        sizes = [(self.bet, self.raise_), (self.bet, self.raise_)]
        return TreeConfig(initial_state=state, starting_pot=self.pot, effective_stack=self.stack,
                          turn_bet_sizes=sizes, river_bet_sizes=sizes, max_raises=self.max_raises)

    def build(self) -> PostFlopGame:
        return PostFlopGame.with_config(self.card_config(), ActionTree(self.tree_config()))

    @property
    def target(self) -> float:
        """0.3% of the starting pot, in chips."""
        return 0.003 * self.pot


def build_all(subgames: Sequence[Subgame], threads: int = 0) -> List[PostFlopGame]:
    """PostFlopGame of every subgame, built on `threads` host threads (0: up
    to 16); the C++ game build runs outside the interpreter lock."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    threads = threads or max(1, min(16, os.cpu_count() or 1))
    if threads == 1 or len(subgames) < 2:
        return [s.build() for s in subgames]
    with ThreadPoolExecutor(max_workers=threads) as pool:
        return list(pool.map(lambda s: s.build(), subgames))


def river_uniform(seed: int = 0) -> Subgame:
    rng = np.random.default_rng(seed)
    board = _board(rng, 5)
    w = np.where(blocked(board), 0.0, 1.0).astype(np.float32)
    return Subgame(board, (w, w.copy()), 100, 100, "50%, 100%, a", "2.5x, a", 1, 1000)


def turn_uniform01(seed: int = 0) -> Subgame:
    rng = np.random.default_rng(seed)
    board = _board(rng, 4)
    bm = blocked(board)
    r0 = np.where(bm, 0.0, rng.random(1326)).astype(np.float32)
    r1 = np.where(bm, 0.0, rng.random(1326)).astype(np.float32)
    return Subgame(board, (r0, r1), 60, 200, "33%, 75%, a", "3x, a", 2, 500)


def river_batch(n: int, seed: int = 0, iterations: int = 300) -> List[Subgame]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        board = _board(rng, 5)
        bm = blocked(board)
        free = np.flatnonzero(~bm)
        ranges = []
        for _p in range(2):
            w = np.zeros(1326, np.float64)
            w[free] = rng.dirichlet(np.full(len(free), 0.5))
            ranges.append(w.astype(np.float32))
        pot = int(round(float(np.exp(rng.uniform(np.log(20.0), np.log(400.0))))))
        spr = float(rng.uniform(0.5, 10.0))
        stack = max(1, int(round(pot * spr)))
        k = int(rng.integers(1, 4))
        pcts = sorted(int(x) for x in rng.choice(BATCH_SIZES, size=k, replace=False))
        cap = int(rng.integers(1, 4))
        menu = _sizes(pcts)
        out.append(Subgame(board, tuple(ranges), pot, stack, menu, menu, cap, iterations))
    return out
