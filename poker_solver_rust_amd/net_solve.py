"""Net-backed depth-limited turn solve (the reference's depth-limited solving
with a value net at the leaves: cfvnet/src/eval/boundary_evaluator.rs:27-111
for the input encoding and the per-river enumeration at turn boundaries,
range-solver game/evaluation.rs:87-288 for the boundary leaves).

Pieces, all on the device except the small encode / combine steps:
  river_training_set   river subgames on a turn board (one random river card
                       each, boundary-like ranges, the turn tree's boundary
                       pots), solved by the batched GPU solver; targets are
                       the per-hand pot-relative values of the solved root
  train_value_net      psr_mlp training on them (BF16 MFMA, Huber, AdamW)
  NetLeaves            per boundary: check-down value plus the net's
                       correction on every river card with the recorded
                       boundary reaches, combined over the rivers
  solve_depth_limited  the turn tree cut at the river deal, DCFR on the GPU
                       with the leaf values refreshed every k iterations

Values.  For hand i of player p at a leaf with opponent reach r, the
pot-relative value is v_i = E[payoff of i] / (pot / 2) against the opponent
range conditioned on no shared card; the leaf's cfvalue is then
v_i * (pot / 2) / num_combinations * R_i with R_i the opponent reach that does
not block hand i -- the device's scaled boundary mode.  Over the river cards
c of a turn boundary the exact mix is v_i = sum_c v_i(c) R_i(c) / sum_c R_i(c)
(R_i(c): opponent reach blocking neither hand i nor c), which follows from
the chance node's 1/44 per card and sum_c R_i(c) = 44 R_i.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence

import numpy as np

from .range_solver import ActionTree, GpuBatchSolver, PostFlopGame, TreeConfig, hand_evaluate
from .value_net import INPUT_DIM, NUM_COMBOS, OUTPUT_DIM, ValueNet
from .workloads import Subgame, _pair_cards, blocked, build_all

# -------------------------------------------------------------- encoding --


def hand_index(game: PostFlopGame, p: int) -> np.ndarray:
    """1326-combo index of every private hand of player p (game order):
    c1 (101 - c1) / 2 + c2 - 1 for c1 < c2 (card.rs:50-57)."""
    pc = np.asarray(game.private_cards(p), np.int64).reshape(-1, 2)
    c1, c2 = pc.min(1), pc.max(1)
    return c1 * (101 - c1) // 2 + c2 - 1


def encode(r0: np.ndarray, r1: np.ndarray, board: Sequence[int], pot: float, stack: float) -> np.ndarray:
    """Input rows: the two 1326 ranges scaled to mean 1 over their non-zero
    combos, the board one-hot, pot / (pot + stack), stack / (pot + stack).
    r0, r1: (n, 1326); board: n boards of 5 cards or one board."""
    r0 = np.atleast_2d(r0).astype(np.float32)
    r1 = np.atleast_2d(r1).astype(np.float32)
    n = r0.shape[0]
    x = np.zeros((n, INPUT_DIM), np.float32)
    for k, r in enumerate((r0, r1)):
        nz = np.maximum((r > 0).sum(1, keepdims=True), 1)
        s = r.sum(1, keepdims=True)
        x[:, k * NUM_COMBOS:(k + 1) * NUM_COMBOS] = np.where(s > 0, r * nz / np.where(s > 0, s, 1), 0)
    boards = np.atleast_2d(np.asarray(board, np.int64))
    if boards.shape[0] == 1 and n > 1:
        boards = np.repeat(boards, n, 0)
    rows = np.repeat(np.arange(n), boards.shape[1])
    x[rows, 2 * NUM_COMBOS + boards.ravel()] = 1.0
    pot = np.broadcast_to(np.asarray(pot, np.float64), (n,))
    stack = np.broadcast_to(np.asarray(stack, np.float64), (n,))
    tot = np.maximum(pot + stack, 1e-9)
    x[:, -2] = pot / tot
    x[:, -1] = stack / tot
    return x


_INCIDENCE = None


def _incidence() -> np.ndarray:
    """(1326, 52) 0/1 matrix: combo holds card."""
    global _INCIDENCE
    if _INCIDENCE is None:
        c1, c2 = _pair_cards()
        m = np.zeros((NUM_COMBOS, 52))
        m[np.arange(NUM_COMBOS), c1] = 1.0
        m[np.arange(NUM_COMBOS), c2] = 1.0
        _INCIDENCE = m
    return _INCIDENCE


def unblocked_mass(r_opp: np.ndarray, extra_card: Optional[int] = None) -> np.ndarray:
    """R[i] for every combo i: opponent weight on combos sharing no card with
    i (and, if given, not holding extra_card).  r_opp: (..., 1326)."""
    c1, c2 = _pair_cards()
    r = np.array(r_opp, np.float64)
    if extra_card is not None:
        r = np.where((c1 == extra_card) | (c2 == extra_card), 0.0, r)
    card = r @ _incidence()
    tot = r.sum(-1, keepdims=True)
    return tot - card[..., c1] - card[..., c2] + r   # the combo itself was removed twice


# ------------------------------------------------------------- training --


@dataclass
class TrainingSet:
    x: np.ndarray
    y: np.ndarray
    mask: np.ndarray
    solve_seconds: float
    build_seconds: float
    solves: int


def boundary_like_ranges(rng: np.random.Generator, bm: np.ndarray) -> np.ndarray:
    """U[0,1] combo weights times 0-3 random 'action probabilities' per combo
    (Beta(0.5, 0.5), mostly near 0 or 1), floored above 0 so every unblocked
    hand stays in the game; blocked combos 0."""
    r = rng.random(NUM_COMBOS)
    for _ in range(int(rng.integers(0, 4))):
        r *= rng.beta(0.5, 0.5, NUM_COMBOS)
    r = np.maximum(r, 1e-4)
    return np.where(bm, 0.0, r).astype(np.float32)


def river_training_set(turn: Subgame, tree_pots: Sequence[int], n: int, seed: int = 0, iterations: int = 300,
                       chunk: int = 4096, log: Callable[[str], None] = print,
                       situations: Optional[list] = None, situation_share: float = 0.5,
                       checkdown: Optional["CheckdownLeaves"] = None) -> TrainingSet:
    """n river subgames on the turn board of `turn` (random river card,
    boundary-like ranges, a pot drawn from `tree_pots` with the stack left
    behind it), solved on the GPU with the reference solve() schedule (stop at
    0.3% of the pot, cap `iterations`).  Targets: pot-relative root values of
    both players per combo; mask 1 on combos present in the game."""
    rng = np.random.default_rng(seed)
    board4 = list(turn.board[:4])
    free = [c for c in range(52) if c not in board4]
    # all-in boundaries (no chips behind) are showdowns, valued exactly without the net
    tree_pots = [p for p in tree_pots if turn.stack - (p - turn.pot) // 2 > 0]
    xs, ys, ms = [], [], []
    t_solve = t_build = 0.0
    for c0 in range(0, n, chunk):
        k = min(chunk, n - c0)
        t = time.time()
        subs = []
        for _ in range(k):
            river = int(rng.choice(free))
            board = board4 + [river]
            bm = blocked(board)
            if situations and rng.random() < situation_share:
                # reaches recorded at a boundary of a depth-limited solve, scaled by
                # a random factor per combo (robustness), on a random river card
                _, reach, pot, stack = situations[int(rng.integers(len(situations)))]
                pot, stack = int(round(pot)), int(round(stack))
                noise = [np.exp(rng.normal(0.0, 0.3, NUM_COMBOS)) for _ in range(2)]
                rr = [np.where(bm, 0.0, np.maximum(reach[p] * noise[p], 1e-4 * reach[p].max())) for p in range(2)]
                if stack <= 0 or min(x.max() for x in rr) <= 0:
                    continue
                # range weights must be <= 1 (range.rs:181-187); values are scale free
                rr = tuple((x / x.max()).astype(np.float32) for x in rr)
                subs.append(Subgame(board, rr, pot, stack, turn.bet, turn.raise_, turn.max_raises, iterations))
                continue
            pot = int(rng.choice(tree_pots))
            stack = max(0, turn.stack - (pot - turn.pot) // 2)
            subs.append(Subgame(board, (boundary_like_ranges(rng, bm), boundary_like_ranges(rng, bm)), pot, stack,
                                turn.bet, turn.raise_, turn.max_raises, iterations))
        k = len(subs)
        games = build_all(subs)
        t_build += time.time() - t
        t = time.time()
        b = GpuBatchSolver(0)
        b.prepare_batch(games, iterations, [s.target for s in subs])
        b.solve(iterations, 5)
        cfv = b.root_cfvalues()
        t_solve += time.time() - t
        y = np.zeros((k, OUTPUT_DIM), np.float32)
        m = np.zeros((k, OUTPUT_DIM), np.float32)
        r0 = np.stack([s.ranges[0] for s in subs]).astype(np.float64)
        r1 = np.stack([s.ranges[1] for s in subs]).astype(np.float64)
        R = (unblocked_mass(r1), unblocked_mass(r0))   # opponent mass per combo of player 0 / 1
        for i, (s, g) in enumerate(zip(subs, games)):
            half = 0.5 * s.pot
            ncomb = g.num_combinations()
            for p in range(2):
                idx = hand_index(g, p)
                rp = R[p][i, idx]
                ok = rp > 0
                y[i, p * NUM_COMBOS + idx[ok]] = cfv[i][p][ok] * ncomb / (half * rp[ok])
                m[i, p * NUM_COMBOS + idx[ok]] = 1.0
        if checkdown is not None:
            # residual targets: exact value minus the check-down value on the same river
            rk = np.array([checkdown.rivers.index(s.board[4]) for s in subs])
            for k in np.unique(rk):
                sel = np.flatnonzero(rk == k)
                base0 = checkdown.values(int(k), r1[sel])
                base1 = checkdown.values(int(k), r0[sel])
                y[sel, :NUM_COMBOS] -= (base0 * m[sel, :NUM_COMBOS]).astype(np.float32)
                y[sel, NUM_COMBOS:] -= (base1 * m[sel, NUM_COMBOS:]).astype(np.float32)
        xs.append(encode(r0, r1, np.array([s.board for s in subs]), np.array([s.pot for s in subs], np.float64),
                         np.array([s.stack for s in subs], np.float64)))
        ys.append(y)
        ms.append(m)
        log(f"training set: {c0 + k}/{n} river solves (build {t_build:.1f}s, solve {t_solve:.1f}s)")
    return TrainingSet(np.concatenate(xs), np.concatenate(ys), np.concatenate(ms), t_solve, t_build, n)


def train_value_net(ts: TrainingSet, epochs: int = 30, batch: int = 512, lr: float = 1e-3, seed: int = 0,
                    weight_decay: float = 1e-4, holdout: float = 0.05, grad_clip_norm: float = 1.0):
    """Trains the default 2706 -> 4 x 512 GELU -> 2652 net, its gradients
    clipped to grad_clip_norm (the reference trainer's default 1.0,
    cfvnet/src/config.rs:284-285; 0: off); returns the net, the per-epoch
    training loss and the hold-out mean absolute error."""
    n = len(ts.x)
    nh = max(1, int(n * holdout))
    net = ValueNet(max_batch=max(batch, 1024), seed=seed)
    net.set_grad_clip(grad_clip_norm)
    net.set_dataset(ts.x[nh:], ts.y[nh:], ts.mask[nh:])
    losses = net.train(epochs, batch, lr=lr, weight_decay=weight_decay, delta=1.0, seed=seed)
    net.train_device_ms = net.last_ms()
    pred = net.forward(ts.x[:nh])
    mae = float((np.abs(pred - ts.y[:nh]) * ts.mask[:nh]).sum() / ts.mask[:nh].sum())
    return net, losses, mae


# ------------------------------------------------------- leaf evaluation --


def river_mix(values: np.ndarray, r_opp: np.ndarray, rivers: Sequence[int]) -> np.ndarray:
    """Combine per-river values (len(rivers), ..., 1326) of one player into
    the turn-boundary value (..., 1326): weights R_i(c) = opponent reach
    (..., 1326) unblocked by hand i and by c; combos holding c get no weight
    from that river."""
    c1, c2 = _pair_cards()
    num = np.zeros(values.shape[1:])
    den = np.zeros(values.shape[1:])
    for k, c in enumerate(rivers):
        w = unblocked_mass(r_opp, c)
        w = np.where((c1 == c) | (c2 == c), 0.0, w)
        num += w * values[k]
        den += w
    mean = values.mean(0)
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), mean)


def zero_sum(v0: np.ndarray, v1: np.ndarray, r0: np.ndarray, r1: np.ndarray):
    """Shift both players' pot-relative values (..., 1326) by a constant each
    so that the leaf is zero-sum: sum_i r0_i R0_i v0_i + sum_j r1_j R1_j v1_j
    = 0 (R: unblocked opponent mass), as exact values of an unraked leaf are;
    half of the violation is taken from each player (the zero-sum output
    layer of DeepStack-style value nets)."""
    R0 = unblocked_mass(r1)
    R1 = unblocked_mass(r0)
    w0 = r0 * R0
    w1 = r1 * R1
    W0 = w0.sum(-1, keepdims=True)
    W1 = w1.sum(-1, keepdims=True)
    e = 0.5 * ((w0 * v0).sum(-1, keepdims=True) + (w1 * v1).sum(-1, keepdims=True))
    d0 = np.where(W0 > 0, e / np.where(W0 > 0, W0, 1.0), 0.0)
    d1 = np.where(W1 > 0, e / np.where(W1 > 0, W1, 1.0), 0.0)
    return v0 - d0, v1 - d1


class NetLeaves:
    """Leaf values from the value net: on every river card the check-down
    value plus the net's correction (the net is trained on exact-minus-
    check-down targets), mixed over the rivers.  Boundaries with no chips
    behind (all-in) are river showdowns: the check-down value is exact there.
    `net` may be a list of nets trained with different seeds: their
    corrections are averaged (an ensemble, which cuts the seed-to-seed spread
    of the error)."""

    def __init__(self, net, checkdown: "CheckdownLeaves"):
        self.nets = list(net) if isinstance(net, (list, tuple)) else [net]
        self.checkdown = checkdown

    def _forward(self, x):
        out = self.nets[0].forward(x)
        for n in self.nets[1:]:
            out = out + n.forward(x)
        return out / len(self.nets)

    def __call__(self, board4, reach, pot, stack):
        return self.values_for([(board4, reach, pot, stack)])[0]

    def values_for(self, items, max_rows: int = 8192):
        """All boundaries of one turn board at once: one net batch over every
        (boundary, river card) row, check-down bases in one vectorised pass."""
        rivers = self.checkdown.rivers
        K = len(rivers)
        r0 = np.stack([np.asarray(it[1][0], np.float64) for it in items])
        r1 = np.stack([np.asarray(it[1][1], np.float64) for it in items])
        base = self.checkdown.per_river_rows(r0, r1)          # (2, K, B, 1326)
        live = [k for k, it in enumerate(items) if it[3] > 0]
        if live:
            c1, c2 = _pair_cards()
            held = np.stack([(c1 == c) | (c2 == c) for c in rivers])               # (K, 1326)
            rows0 = np.where(held[None], 0.0, r0[live][:, None]).reshape(-1, NUM_COMBOS)
            rows1 = np.where(held[None], 0.0, r1[live][:, None]).reshape(-1, NUM_COMBOS)
            board4 = list(items[live[0]][0])
            boards = np.array([board4 + [c] for c in rivers] * len(live))
            pots = np.repeat([items[k][2] for k in live], K)
            stacks = np.repeat([items[k][3] for k in live], K)
            x = encode(rows0, rows1, boards, pots, stacks)
            out = np.concatenate([self._forward(x[i:i + max_rows]) for i in range(0, len(x), max_rows)])
            out = out.reshape(len(live), K, OUTPUT_DIM).transpose(1, 0, 2)      # (K, live, 2652)
            v0 = base[0][:, live] + out[:, :, :NUM_COMBOS]
            v1 = base[1][:, live] + out[:, :, NUM_COMBOS:]
            v0, v1 = zero_sum(v0, v1, rows0.reshape(len(live), K, -1).transpose(1, 0, 2),
                              rows1.reshape(len(live), K, -1).transpose(1, 0, 2))
            base[0][:, live] = v0
            base[1][:, live] = v1
        m0 = river_mix(base[0], r1, rivers)
        m1 = river_mix(base[1], r0, rivers)
        return [(m0[k], m1[k]) for k in range(len(items))]


class CheckdownLeaves:
    """Leaf values without a net: both players check the river down --
    showdown equity over every river card, v_i(c) = (win - lose mass) / R_i(c),
    mixed over the rivers like the net's values.  Vectorised per river card:
    the strength order is fixed, so the mass below / above hand i is a prefix
    sum, and the card-removal terms only involve the 51 combos holding each of
    i's cards (a 51 x 51 comparison per card)."""

    def __init__(self, board4: Sequence[int]):
        self.board4 = list(board4)
        self.rivers = [c for c in range(52) if c not in board4]
        c1, c2 = _pair_cards()
        self.tables = []
        for c in self.rivers:
            board = self.board4 + [c]
            bm = blocked(board)
            s = np.zeros(NUM_COMBOS, np.int64)
            for i in np.flatnonzero(~bm):
                s[i] = hand_evaluate(board + [int(c1[i]), int(c2[i])]) + 1
            order = np.argsort(s, kind="stable")
            ss = s[order]
            lo = np.searchsorted(ss, s, side="left")    # combos strictly weaker (incl. invalid, weight 0)
            hi = np.searchsorted(ss, s, side="right")   # combos not stronger
            cards = []
            for card in range(52):
                members = np.flatnonzero(((c1 == card) | (c2 == card)) & ~bm)
                sm = s[members]
                cards.append((members, (sm[None, :] < sm[:, None]).astype(np.float64),
                              (sm[None, :] > sm[:, None]).astype(np.float64)))
            self.tables.append((s > 0, order, lo, hi, cards))

    def values(self, k: int, r_opp: np.ndarray) -> np.ndarray:
        """(B, 1326) check-down values on river self.rivers[k] against
        opponent reach rows r_opp (B, 1326)."""
        valid, order, lo, hi, cards = self.tables[k]
        r = np.where(valid, np.atleast_2d(np.asarray(r_opp, np.float64)), 0.0)
        B = r.shape[0]
        pref = np.concatenate([np.zeros((B, 1)), np.cumsum(r[:, order], axis=1)], axis=1)
        tot = pref[:, -1:]
        win = pref[:, lo]
        lose = tot - pref[:, hi]
        c1, c2 = _pair_cards()
        for card, (members, lower, upper) in enumerate(cards):
            if len(members) == 0:
                continue
            rm = r[:, members]
            below = rm @ lower.T     # (B, m): mass of combos with `card` weaker than each member
            above = rm @ upper.T
            # subtract, for every combo holding `card`, the part of its totals that shares the card
            win[:, members] -= below
            lose[:, members] -= above
        R = unblocked_mass(r)
        out = np.where(valid & (R > 0), (win - lose) / np.where(R > 0, R, 1.0), 0.0)
        return out

    def per_river_rows(self, r0: np.ndarray, r1: np.ndarray) -> np.ndarray:
        """(2, K, B, 1326): check-down values of both players on every river
        card for B boundaries at once (reach rows r0, r1: (B, 1326))."""
        c1, c2 = _pair_cards()
        r0 = np.atleast_2d(r0)
        r1 = np.atleast_2d(r1)
        out = np.zeros((2, len(self.rivers), r0.shape[0], NUM_COMBOS))
        for k, c in enumerate(self.rivers):
            held = (c1 == c) | (c2 == c)
            out[0, k] = self.values(k, np.where(held, 0.0, r1))
            out[1, k] = self.values(k, np.where(held, 0.0, r0))
        return out

    def per_river(self, reach: List[np.ndarray]):
        """(2, K, 1326): check-down values of both players on every river."""
        return self.per_river_rows(reach[0], reach[1])[:, :, 0]

    def values_for(self, items):
        r0 = np.stack([np.asarray(it[1][0], np.float64) for it in items])
        r1 = np.stack([np.asarray(it[1][1], np.float64) for it in items])
        v = self.per_river_rows(r0, r1)
        m0 = river_mix(v[0], r1, self.rivers)
        m1 = river_mix(v[1], r0, self.rivers)
        return [(m0[k], m1[k]) for k in range(len(items))]

    def __call__(self, board4, reach, pot, stack):
        return self.values_for([(board4, reach, pot, stack)])[0]


class DeviceCheckdownLeaves:
    """The check-down values of CheckdownLeaves computed by the device
    showdown evaluator (game/evaluation.rs:357-425): for each of `slots`
    boundaries and each river card a river subgame whose tree is check /
    check, built once with every unblocked combo in both ranges.  A refresh
    replaces their initial weights by the boundary reaches (river-card
    combos dropped) and runs one cfvalue pass over all of them; hand i's
    root cfvalue is (half pot / num_combinations) (win - lose mass), so
    v_i(c) = cfv_i num_combinations / (half pot R_i(c)).  Same interface as
    CheckdownLeaves (per_river_rows, values_for, values)."""

    def __init__(self, board4: Sequence[int], slots: int, pot: int = 100, stack: int = 100):
        from .range_solver import BoardState, CardConfig, Range
        from .workloads import _pair_cards as _pc

        self.board4 = list(board4)
        self.rivers = [c for c in range(52) if c not in board4]
        self.slots = slots
        self.half = 0.5 * pot
        tc = TreeConfig(initial_state=BoardState.River, starting_pot=pot, effective_stack=stack)
        tree = ActionTree(tc)
        c1, c2 = _pc()
        self.games, self.idx, self.ncomb = [], [], []
        for c in self.rivers:
            board = self.board4 + [c]
            w = np.where(blocked(board), 0.0, 1.0).astype(np.float32)
            cc = CardConfig(range=(Range(w), Range(w)), flop=sorted(board[:3]), turn=board[3], river=c)
            g = PostFlopGame.with_config(cc, tree)
            self.games.append(g)
            self.idx.append(hand_index(g, 0))   # both ranges hold the same combos
            self.ncomb.append(g.num_combinations())
        self.held = np.stack([(c1 == c) | (c2 == c) for c in self.rivers])
        self.solver = GpuBatchSolver(0)
        # game (slot k, river j) at k * K + j; the check / check tree needs no iterations
        self.solver.prepare_batch([g for _ in range(slots) for g in self.games], 1)

    def per_river_rows(self, r0: np.ndarray, r1: np.ndarray) -> np.ndarray:
        """(2, K, B, 1326) check-down values of both players on every river
        for B <= slots boundaries (reach rows r0, r1: (B, 1326))."""
        r0 = np.atleast_2d(np.asarray(r0, np.float64))
        r1 = np.atleast_2d(np.asarray(r1, np.float64))
        B, K = r0.shape[0], len(self.rivers)
        if B > self.slots:
            raise ValueError(f"{B} boundaries, {self.slots} slots")
        # weights: every slot gets a row (unused slots repeat row 0); scale free, so
        # normalised per row to the [0, 1] a range weight must lie in
        w, tops = [], []
        for r in (r0, r1):
            rows = np.concatenate([r, np.repeat(r[:1], self.slots - B, 0)]) if B < self.slots else r
            top = rows.max(1, keepdims=True)
            tops.append(top[:B, 0])
            rows = np.where(top > 0, rows / np.where(top > 0, top, 1.0), 0.0)
            w.append(np.concatenate([rows[k][self.idx[j]] for k in range(self.slots) for j in range(K)]))
        self.solver.set_initial_weights(w[0].astype(np.float32), w[1].astype(np.float32))
        self.solver.compute_root_cfvalues()
        cf = [self.solver.root_cfvalues_all(p) for p in range(2)]
        out = np.zeros((2, K, B, NUM_COMBOS))
        n = len(self.idx[0])
        for p, r_opp in ((0, r1), (1, r0)):
            top = np.maximum(tops[p ^ 1], 1e-300)     # the opponent reach on the device is r_opp / top
            for j in range(K):
                ro = np.where(self.held[j], 0.0, r_opp)                         # (B, 1326)
                R = unblocked_mass(ro)[:, self.idx[j]] / top[:, None]           # (B, H), device scale
                for k in range(B):
                    v = cf[p][(k * K + j) * n:(k * K + j + 1) * n].astype(np.float64)
                    ok = R[k] > 0
                    out[p, j, k, self.idx[j]] = np.where(
                        ok, v * self.ncomb[j] / (self.half * np.where(ok, R[k], 1.0)), 0.0)
        return out

    def per_river(self, reach: List[np.ndarray]):
        return self.per_river_rows(reach[0], reach[1])[:, :, 0]

    def values_for(self, items):
        r0 = np.stack([np.asarray(it[1][0], np.float64) for it in items])
        r1 = np.stack([np.asarray(it[1][1], np.float64) for it in items])
        v = self.per_river_rows(r0, r1)
        m0 = river_mix(v[0], r1, self.rivers)
        m1 = river_mix(v[1], r0, self.rivers)
        return [(m0[k], m1[k]) for k in range(len(items))]

    def __call__(self, board4, reach, pot, stack):
        return self.values_for([(board4, reach, pot, stack)])[0]


class ExactLeaves:
    """Leaf values from exact river solves: every (boundary, river card)
    subgame with the recorded reaches as ranges, solved in one batched GPU
    launch per refresh (stop at 0.3% of the pot, cap `iterations`) -- the
    fallback and the yardstick of the net."""

    def __init__(self, turn: Subgame, iterations: int = 300):
        self.turn = turn
        self.iterations = iterations
        self.checkdown = CheckdownLeaves(turn.board[:4])

    def values_for(self, items):
        """items: list of (board4, reach, pot, stack); returns [(v0, v1)]."""
        c1, c2 = _pair_cards()
        out = [None] * len(items)
        subs, where = [], []
        for k, (board4, reach, pot, stack) in enumerate(items):
            if stack <= 0:
                out[k] = self.checkdown(board4, reach, pot, stack)
                continue
            rivers = [c for c in range(52) if c not in board4]
            for c in rivers:
                held = (c1 == c) | (c2 == c)
                r0 = np.where(held, 0.0, reach[0])
                r1 = np.where(held, 0.0, reach[1])
                if r0.max() <= 0 or r1.max() <= 0:
                    continue
                r0 = (r0 / r0.max()).astype(np.float32)   # range weights <= 1; values are scale free
                r1 = (r1 / r1.max()).astype(np.float32)
                subs.append(Subgame(list(board4) + [c], (r0, r1), int(round(pot)), int(round(stack)), self.turn.bet,
                                    self.turn.raise_, self.turn.max_raises, self.iterations))
                where.append((k, c))
        vals = {}
        if subs:
            games = build_all(subs)
            b = GpuBatchSolver(0)
            b.prepare_batch(games, self.iterations, [s.target for s in subs])
            b.solve(self.iterations, 5)
            cfv = b.root_cfvalues()
            for i, (s, g) in enumerate(zip(subs, games)):
                half = 0.5 * s.pot
                ncomb = g.num_combinations()
                v = []
                for p in range(2):
                    full = np.zeros(NUM_COMBOS)
                    idx = hand_index(g, p)
                    R = unblocked_mass(s.ranges[p ^ 1].astype(np.float64))[idx]
                    ok = R > 0
                    full[idx[ok]] = cfv[i][p][ok] * ncomb / (half * R[ok])
                    v.append(full)
                vals[where[i]] = v
        for k, (board4, reach, pot, stack) in enumerate(items):
            if out[k] is not None:
                continue
            rivers = [c for c in range(52) if c not in board4]
            v0 = np.stack([vals.get((k, c), [np.zeros(NUM_COMBOS)] * 2)[0] for c in rivers])
            v1 = np.stack([vals.get((k, c), [np.zeros(NUM_COMBOS)] * 2)[1] for c in rivers])
            out[k] = (river_mix(v0, reach[1], rivers), river_mix(v1, reach[0], rivers))
        return out


# ------------------------------------------------------ turn solve --


def depth_limited_tree(turn: Subgame) -> TreeConfig:
    tc = turn.tree_config()
    tc.depth_limit = 0
    return tc


def boundary_stacks(turn: Subgame, game: PostFlopGame) -> List[float]:
    return [max(0.0, turn.stack - (game.boundary_pot(b) - turn.pot) / 2.0) for b in range(game.num_boundary_nodes())]


def solve_depth_limited(turn: Subgame, leaf_fn, iterations: int = 500, refresh: int = 10,
                        log: Callable[[str], None] = print, snapshots: Optional[list] = None):
    """DCFR on the turn tree cut at the river deal; every `refresh`
    iterations the leaves get leaf_fn(board4, [reach0, reach1], pot, stack)
    (1326-combo vectors) from the reaches recorded at each boundary (or, if
    leaf_fn has values_for, all boundaries at once).  `snapshots`, if a list,
    receives the (board4, reach, pot, stack) items of every refresh.
    Returns (game, solver, seconds in leaf evaluation, seconds in DCFR)."""
    game = PostFlopGame.with_config(turn.card_config(), ActionTree(depth_limited_tree(turn)))
    nb = game.num_boundary_nodes()
    idx = [hand_index(game, p) for p in range(2)]
    w = [game.initial_weights(p) for p in range(2)]
    stacks = boundary_stacks(turn, game)
    b = GpuBatchSolver(0)
    b.prepare_batch([game], iterations)
    t_leaf = t_dcfr = 0.0
    for t0 in range(0, iterations, refresh):
        t = time.time()
        items = []
        for k in range(nb):
            reach = []
            for p in range(2):
                r = b.boundary_reach(0, k, p) if t0 > 0 else w[p]
                full = np.zeros(NUM_COMBOS)
                full[idx[p]] = r
                reach.append(full)
            items.append((turn.board[:4], reach, float(game.boundary_pot(k)), stacks[k]))
        if snapshots is not None:
            snapshots.append(items)
        if hasattr(leaf_fn, "values_for"):
            vals = leaf_fn.values_for(items)
        else:
            vals = [leaf_fn(*it) for it in items]
        for k, v in enumerate(vals):
            for p in range(2):
                b.set_boundary_cfvs(0, k, p, v[p][idx[p]].astype(np.float32), raw=False)
        t_leaf += time.time() - t
        t = time.time()
        b.run_iterations(t0, min(t0 + refresh, iterations))
        t_dcfr += time.time() - t
        if (t0 // refresh) % 10 == 0:
            log(f"depth-limited solve: {min(t0 + refresh, iterations)}/{iterations} iterations")
    # The solution is the average strategy: a cfvalue pass records the reaches
    # it produces at every boundary, and the leaves are valued for those (the
    # current strategy's reaches, used above, do not converge under CFR).
    t = time.time()
    for _ in range(2):
        b.compute_root_cfvalues()
        refresh_leaves(turn, game, b, leaf_fn, idx, stacks)
    t_leaf += time.time() - t
    return game, b, t_leaf, t_dcfr


def refresh_leaves(turn, game, b, leaf_fn, idx, stacks, snapshots=None):
    """Values every boundary of game 0 of solver b from its recorded reaches."""
    items = []
    for k in range(game.num_boundary_nodes()):
        reach = []
        for p in range(2):
            full = np.zeros(NUM_COMBOS)
            full[idx[p]] = b.boundary_reach(0, k, p)
            reach.append(full)
        items.append((turn.board[:4], reach, float(game.boundary_pot(k)), stacks[k]))
    if snapshots is not None:
        snapshots.append(items)
    vals = leaf_fn.values_for(items) if hasattr(leaf_fn, "values_for") else [leaf_fn(*it) for it in items]
    for k, v in enumerate(vals):
        for p in range(2):
            b.set_boundary_cfvs(0, k, p, v[p][idx[p]].astype(np.float32), raw=False)


def boundary_items(turn: Subgame, game: PostFlopGame, b: GpuBatchSolver) -> list:
    """(board4, [reach0, reach1], pot, stack) of every boundary of game 0 of b,
    from the reaches it recorded last (after compute_root_cfvalues: those of
    the average strategy)."""
    idx = [hand_index(game, p) for p in range(2)]
    stacks = boundary_stacks(turn, game)
    items = []
    for k in range(game.num_boundary_nodes()):
        reach = []
        for p in range(2):
            full = np.zeros(NUM_COMBOS)
            full[idx[p]] = b.boundary_reach(0, k, p)
            reach.append(full)
        items.append((turn.board[:4], reach, float(game.boundary_pot(k)), stacks[k]))
    return items


def leaf_error(items, got, ref) -> float:
    """Mean absolute leaf-value error as a fraction of the pot, weighted by
    how much each hand weighs at its leaf (own reach times unblocked opponent
    reach), over both players and every leaf: sum |v - v_ref| w / 2 / sum w
    (v are values per half pot)."""
    num = den = 0.0
    for it, g, r in zip(items, got, ref):
        for p in range(2):
            w = it[1][p] * unblocked_mass(it[1][p ^ 1])
            num += float((np.abs(g[p] - r[p]) * w).sum()) * 0.5
            den += float(w.sum())
    return num / max(den, 1e-30)


def root_ev_error(exact_b: GpuBatchSolver, other_b: GpuBatchSolver, game_exact: PostFlopGame, pot: float) -> dict:
    """Root expected values (query.rs:639-666) of both players, as a fraction
    of the pot: weighted mean and max absolute difference, and the difference
    of the game value (range-weighted mean EV)."""
    out = {}
    for p in range(2):
        e = exact_b.expected_values(0, p).astype(np.float64)
        o = other_b.expected_values(0, p).astype(np.float64)
        nw = exact_b.normalized_weights(0, p).astype(np.float64)
        d = np.abs(e - o)
        out[f"p{p}_mean_abs_pot"] = float((nw * d).sum() / nw.sum() / pot)
        out[f"p{p}_max_abs_pot"] = float(d.max() / pot)
        out[f"p{p}_value_diff_pot"] = float(abs((nw * (e - o)).sum()) / nw.sum() / pot)
    return out
