"""ctypes mirror of the reference range-solver API over libpsr_dcfr.so.

Reference map (crates/range-solver/src unless noted):
  card_pair_to_index / index_to_card_pair   card.rs:50-65
  Range (from_str)                           range.rs:721-760
  TreeConfig / ActionTree                    action_tree.rs:130-334
  CardConfig                                 card.rs:20-32
  PostFlopGame.with_config                   game/interpreter.rs:128-166
  GpuSolverConfig / GpuSolveResult / gpu_solve_hand_parallel
                                             gpu-range-solver lib.rs:14-87
  GpuBatchSolver                             gpu-range-solver batch.rs:79-610
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

NOT_DEALT = 0xFF
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class PsrError(RuntimeError):
    """Error reported by the library (the reference returns Err / panics)."""


def lib_path() -> str:
    return os.path.join(_HERE, "libpsr_dcfr.so")


class _TreeConfigC(C.Structure):
    _fields_ = [
        ("initial_state", C.c_int32),
        ("starting_pot", C.c_int32),
        ("effective_stack", C.c_int32),
        ("rake_rate", C.c_double),
        ("rake_cap", C.c_double),
        ("bet", (C.c_char_p * 2) * 3),
        ("raise_", (C.c_char_p * 2) * 3),
        ("donk", C.c_char_p * 2),
        ("add_allin_threshold", C.c_double),
        ("force_allin_threshold", C.c_double),
        ("merging_threshold", C.c_double),
        ("depth_limit", C.c_int32),
        ("max_raises", C.c_int32),
    ]


class _ActionC(C.Structure):
    _fields_ = [("kind", C.c_int32), ("amount", C.c_int32)]


class _CardConfigC(C.Structure):
    _fields_ = [
        ("range", C.POINTER(C.c_float) * 2),
        ("flop", C.c_uint8 * 3),
        ("turn", C.c_uint8),
        ("river", C.c_uint8),
    ]


class _SolverConfigC(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_uint32),
        ("target_exploitability", C.c_float),
        ("print_progress", C.c_int32),
        ("check_interval", C.c_int32),
    ]


class _SolveResultC(C.Structure):
    _fields_ = [("exploitability", C.c_float), ("iterations_run", C.c_uint32)]


def load_library(path: Optional[str] = None):
    """Loads libpsr_dcfr.so; raises if it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or lib_path()
    if not os.path.exists(p):
        raise PsrError(f"{p} is missing: build it with `python -m poker_solver_rust_amd.build`")
    lib = C.CDLL(p)
    vp = C.c_void_p
    sig = {
        "psr_last_error": (C.c_char_p, []),
        "psr_card_pair_to_index": (C.c_int, [C.c_uint8, C.c_uint8]),
        "psr_index_to_card_pair": (C.c_int, [C.c_int, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]),
        "psr_card_from_str": (C.c_int, [C.c_char_p]),
        "psr_range_parse": (C.c_int, [C.c_char_p, C.POINTER(C.c_float)]),
        "psr_hand_evaluate": (C.c_int, [C.POINTER(C.c_uint8), C.c_int]),
        "psr_action_tree_build": (C.c_int, [C.POINTER(_TreeConfigC), C.POINTER(vp)]),
        "psr_action_tree_free": (None, [vp]),
        "psr_action_tree_actions": (C.c_int, [vp, C.POINTER(_ActionC), C.c_int, C.POINTER(_ActionC), C.c_int]),
        "psr_action_tree_count_nodes": (C.c_int, [vp, C.POINTER(C.c_uint64)]),
        "psr_action_tree_node_info": (C.c_int, [vp, C.POINTER(_ActionC), C.c_int, C.POINTER(C.c_int32)]),
        "psr_action_tree_invalid_terminals": (C.c_int, [vp]),
        "psr_game_build": (C.c_int, [C.POINTER(_CardConfigC), vp, C.POINTER(vp)]),
        "psr_game_build_ex": (C.c_int, [C.POINTER(_CardConfigC), vp, C.c_uint32, C.POINTER(vp)]),
        "psr_game_free": (None, [vp]),
        "psr_game_num_hands": (C.c_int, [vp, C.c_int]),
        "psr_game_private_cards": (C.c_int, [vp, C.c_int, C.POINTER(C.c_uint8)]),
        "psr_game_initial_weights": (C.c_int, [vp, C.c_int, C.POINTER(C.c_float)]),
        "psr_game_num_nodes": (C.c_int64, [vp]),
        "psr_game_num_combinations": (C.c_double, [vp]),
        "psr_game_num_storage": (C.c_int64, [vp]),
        "psr_game_serialize": (C.c_int64, [vp, C.POINTER(C.c_uint8), C.c_int64]),
        "psr_gpu_solve": (C.c_int, [vp, C.POINTER(_SolverConfigC), C.POINTER(_SolveResultC), C.POINTER(C.c_float), C.c_int64]),
        "psr_batch_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "psr_batch_destroy": (None, [vp]),
        "psr_batch_prepare": (C.c_int, [vp, C.POINTER(vp), C.c_int, C.c_uint32, C.POINTER(C.c_float)]),
        "psr_batch_reset": (C.c_int, [vp]),
        "psr_batch_set_targets": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "psr_batch_run_iterations": (C.c_int, [vp, C.c_uint32, C.c_uint32]),
        "psr_batch_solve": (C.c_int, [vp, C.c_uint32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_uint32)]),
        "psr_batch_exploitability": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "psr_batch_compute_root_cfvalues": (C.c_int, [vp]),
        "psr_batch_root_cfvalues": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_float)]),
        "psr_batch_root_cfvalues_all": (C.c_int, [vp, C.c_int, C.POINTER(C.c_float), C.c_int64]),
        "psr_batch_set_initial_weights": (C.c_int, [vp, C.c_int, C.POINTER(C.c_float), C.c_int64,
                                                    C.POINTER(C.c_float), C.c_int64]),
        "psr_batch_strategy_sum": (C.c_int, [vp, C.c_int, C.POINTER(C.c_float), C.c_int64]),
        "psr_batch_regrets": (C.c_int, [vp, C.c_int, C.POINTER(C.c_float), C.c_int64]),
        "psr_batch_last_kernel_ms": (C.c_float, [vp]),
        "psr_batch_reach_at_nodes": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_int64), C.c_int,
                                     C.POINTER(C.c_float)]),
        "psr_batch_launch_info": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "psr_batch_set_launch": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32]),
        "psr_batch_autotune": (C.c_int, [vp, C.c_uint32, C.POINTER(C.c_int32)]),
        "psr_batch_set_depth_first": (C.c_int, [vp, C.c_int32]),
        "psr_batch_depth_first_info": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                                 C.POINTER(C.c_int64)]),
        "psr_game_num_boundaries": (C.c_int, [vp]),
        "psr_game_boundary_node": (C.c_int64, [vp, C.c_int]),
        "psr_game_boundary_pot": (C.c_int, [vp, C.c_int]),
        "psr_batch_num_boundaries": (C.c_int, [vp, C.c_int]),
        "psr_batch_set_boundary_cfvs": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int,
                                                  C.c_int]),
        "psr_batch_update_leaf_cfvs": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float),
                                                 C.c_int64]),
        "psr_batch_boundary_reach": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int]),
        "psr_batch_normalized_weights": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int]),
        "psr_game_topology": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int64]),
        "psr_batch_set_leaf_injection": (C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int,
                                                   C.c_int32]),
        "psr_batch_download_reach": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int64, C.c_int32]),
        "psr_batch_strategy_sum_edges": (C.c_int, [vp, C.c_int, C.POINTER(C.c_float), C.c_int64, C.c_int32]),
        "psr_mlp_create": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]),
        "psr_mlp_destroy": (None, [vp]),
        "psr_mlp_num_params": (C.c_int64, [vp]),
        "psr_mlp_get_params": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int64]),
        "psr_mlp_set_params": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int64]),
        "psr_mlp_forward": (C.c_int, [vp, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)]),
        "psr_mlp_gradients": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int,
                                        C.c_float, C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_int64]),
        "psr_mlp_set_dataset": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                          C.c_int]),
        "psr_mlp_train": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_uint64,
                                    C.POINTER(C.c_double)]),
        "psr_mlp_last_ms": (C.c_float, [vp]),
        "psr_mlp_set_grad_clip": (C.c_int, [vp, C.c_float]),
        "psr_mlp_last_grad_norm": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "psr_batch_expected_values_detail": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int64]),
        "psr_batch_expected_values": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int]),
        "psr_layout_validate": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "psr_batch_set_storage": (C.c_int, [vp, C.c_int]),
        "psr_batch_exploitability_f64": (C.c_int, [vp, C.POINTER(C.c_double)]),
        "psr_batch_raw_state": (C.c_int, [vp, C.c_int, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16),
                                          C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64, C.c_int32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def _check(rc):
    if rc is not None and rc < 0:
        raise PsrError(load_library().psr_last_error().decode())
    return rc


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# ------------------------------------------------------------------ cards --

def card_pair_to_index(c1: int, c2: int) -> int:
    return _check(load_library().psr_card_pair_to_index(c1, c2))


def index_to_card_pair(index: int) -> Tuple[int, int]:
    a, b = C.c_uint8(), C.c_uint8()
    _check(load_library().psr_index_to_card_pair(index, C.byref(a), C.byref(b)))
    return a.value, b.value


def card_from_str(s: str) -> int:
    return _check(load_library().psr_card_from_str(s.encode()))


def flop_from_str(s: str) -> List[int]:
    """card.rs flop_from_str: 3 cards, optionally space separated, sorted."""
    s = s.replace(" ", "")
    if len(s) != 6:
        raise PsrError("Expected exactly three cards")
    cards = sorted(card_from_str(s[i:i + 2]) for i in range(0, 6, 2))
    if cards[0] == cards[1] or cards[1] == cards[2]:
        raise PsrError("Cards must be unique")
    return cards


def hand_evaluate(cards: Sequence[int]) -> int:
    arr = (C.c_uint8 * len(cards))(*cards)
    return _check(load_library().psr_hand_evaluate(arr, len(cards)))


class Range:
    """1326 weights in card_pair_to_index order (range.rs:38-40)."""

    def __init__(self, data=None):
        self.data = np.zeros(1326, np.float32) if data is None else np.asarray(data, np.float32).copy()
        if self.data.shape != (1326,):
            raise PsrError("Expected exactly 1326 elements")

    @classmethod
    def parse(cls, s: str) -> "Range":
        out = np.zeros(1326, np.float32)
        _check(load_library().psr_range_parse(s.encode(), _fptr(out)))
        return cls(out)

    from_str = parse

    @classmethod
    def ones(cls) -> "Range":
        return cls(np.ones(1326, np.float32))

    def get_weight_by_cards(self, c1: int, c2: int) -> float:
        return float(self.data[card_pair_to_index(c1, c2)])


# ------------------------------------------------------------------- tree --

class BoardState(enum.IntEnum):
    """action_tree.rs:78-100: Flop < Turn < River, default Flop."""
    Flop = 0
    Turn = 1
    River = 2

    def __str__(self):
        return self.name


class Action(enum.IntEnum):
    """Action kinds of action_tree.rs:37-51 in their Ord order.  An action is
    the pair (kind, amount); pairs compare as the reference's derived Ord
    (variant first, then amount) and hash by value."""
    None_ = 0
    Fold = 1
    Check = 2
    Call = 3
    Bet = 4
    Raise = 5
    AllIn = 6
    Chance = 7


# action_tree.rs:10-33 player field flags
PLAYER_OOP = 0
PLAYER_IP = 1
PLAYER_CHANCE = 2
PLAYER_MASK = 3
PLAYER_CHANCE_FLAG = 4
PLAYER_TERMINAL_FLAG = 8
PLAYER_FOLD_FLAG = 24
PLAYER_DEPTH_BOUNDARY_FLAG = 40

_RANKS, _SUITS = "23456789TJQKA", "cdhs"


def card_to_string(card: int) -> str:
    """card.rs card_to_string: "2c" .. "As"; raises for cards outside 0..51."""
    if not 0 <= int(card) < 52:
        raise PsrError(f"Invalid card: {card}")
    return _RANKS[card >> 2] + _SUITS[card & 3]


def action_to_string(action: Tuple[int, int]) -> str:
    """action_tree.rs:53-72 Display for Action."""
    kind, amount = Action(action[0]), int(action[1])
    if kind in (Action.None_, Action.Fold, Action.Check, Action.Call):
        return "None" if kind == Action.None_ else kind.name
    if kind == Action.Chance:
        return f"Chance({card_to_string(amount)})" if 0 <= amount < 52 else f"Chance({amount})"
    return f"{kind.name}({amount})"


def _act(kind, amount=0):
    return (int(kind), int(amount))


@dataclass
class TreeConfig:
    """action_tree.rs:130-192.  Bet sizes are (bet, raise) strings per player
    and street, e.g. river_bet_sizes=[("50%, a", "2.5x"), ("50%, a", "2.5x")]."""
    initial_state: BoardState = BoardState.Flop
    starting_pot: int = 0
    effective_stack: int = 0
    rake_rate: float = 0.0
    rake_cap: float = 0.0
    flop_bet_sizes: Sequence[Tuple[str, str]] = (("", ""), ("", ""))
    turn_bet_sizes: Sequence[Tuple[str, str]] = (("", ""), ("", ""))
    river_bet_sizes: Sequence[Tuple[str, str]] = (("", ""), ("", ""))
    turn_donk_sizes: Optional[str] = None
    river_donk_sizes: Optional[str] = None
    add_allin_threshold: float = 0.0
    force_allin_threshold: float = 0.0
    merging_threshold: float = 0.0
    depth_limit: Optional[int] = None
    max_raises: int = 0   # extension: 0 = unlimited (reference behaviour)

    def _to_c(self):
        c = _TreeConfigC()
        c.initial_state = int(self.initial_state)
        c.starting_pot = self.starting_pot
        c.effective_stack = self.effective_stack
        c.rake_rate = self.rake_rate
        c.rake_cap = self.rake_cap
        keep = []
        for s, sizes in enumerate((self.flop_bet_sizes, self.turn_bet_sizes, self.river_bet_sizes)):
            for p in range(2):
                b, r = sizes[p]
                bb, rb = b.encode(), r.encode()
                keep += [bb, rb]
                c.bet[s][p] = bb
                c.raise_[s][p] = rb
        for i, d in enumerate((self.turn_donk_sizes, self.river_donk_sizes)):
            if d is not None:
                db = d.encode()
                keep.append(db)
                c.donk[i] = db
        c.add_allin_threshold = self.add_allin_threshold
        c.force_allin_threshold = self.force_allin_threshold
        c.merging_threshold = self.merging_threshold
        c.depth_limit = -1 if self.depth_limit is None else int(self.depth_limit)
        c.max_raises = int(self.max_raises)
        return c, keep


def _history_c(history):
    return (_ActionC * max(len(history), 1))(*[_ActionC(int(k), int(a)) for k, a in history])


class ActionTree:
    """action_tree.rs:326-334 ActionTree::new, its navigation
    (play / undo / back_to_root / apply_history / history,
    action_tree.rs:416-481) and query helpers.  Actions are (kind, amount)
    pairs; chance actions are left out of a history, as in the reference."""

    def __init__(self, config: TreeConfig):
        lib = load_library()
        c, keep = config._to_c()
        h = C.c_void_p()
        _check(lib.psr_action_tree_build(C.byref(c), C.byref(h)))
        self._h = h
        self.config = config
        self._history: List[Tuple[Action, int]] = []

    def __del__(self):
        if getattr(self, "_h", None):
            load_library().psr_action_tree_free(self._h)
            self._h = None

    def available_actions(self, history: Optional[Sequence[Tuple[int, int]]] = None) -> List[Tuple[Action, int]]:
        """Actions at the node `history` reaches (the current node when None),
        after any chance node there (action_tree.rs:452-455)."""
        lib = load_library()
        history = self._history if history is None else history
        out = (_ActionC * 64)()
        n = _check(lib.psr_action_tree_actions(self._h, _history_c(history), len(history), out, 64))
        return [(Action(out[i].kind), out[i].amount) for i in range(n)]

    def _node_info(self, history):
        info = (C.c_int32 * 4)()
        _check(load_library().psr_action_tree_node_info(self._h, _history_c(history), len(history), info))
        return list(info)

    def history(self) -> List[Tuple[Action, int]]:
        return list(self._history)

    def play(self, action: Tuple[int, int]):
        """action_tree.rs:462-470: the action must be available here."""
        action = (Action(action[0]), int(action[1]))
        if action not in self.available_actions():
            raise PsrError(f"Action `{action_to_string(action)}` is not available")
        self._history.append(action)

    def undo(self):
        if not self._history:
            raise PsrError("No action to undo")
        self._history.pop()

    def back_to_root(self):
        self._history = []

    def apply_history(self, history: Sequence[Tuple[int, int]]):
        self.back_to_root()
        for a in history:
            self.play(a)

    def is_terminal_node(self) -> bool:
        return bool(self._node_info(self._history)[0])

    def is_chance_node(self) -> bool:
        return bool(self._node_info(self._history)[1])

    def total_bet_amount(self) -> List[int]:
        """action_tree.rs:523-527: chips each player has put in along the
        current history."""
        return self._node_info(self._history)[2:4]

    def num_invalid_terminals(self) -> int:
        """len(invalid_terminals()), action_tree.rs:356-362."""
        return _check(load_library().psr_action_tree_invalid_terminals(self._h))

    def count_nodes(self) -> List[int]:
        out = (C.c_uint64 * 3)()
        _check(load_library().psr_action_tree_count_nodes(self._h, out))
        return list(out)


@dataclass
class CardConfig:
    """card.rs:20-32."""
    range: Tuple[Range, Range] = field(default_factory=lambda: (Range(), Range()))
    flop: Sequence[int] = (NOT_DEALT, NOT_DEALT, NOT_DEALT)
    turn: int = NOT_DEALT
    river: int = NOT_DEALT


class PostFlopGame:
    """game/interpreter.rs:128-166 PostFlopGame::with_config."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def with_config(cls, card_config: CardConfig, action_tree: ActionTree, isomorphism: bool = True) -> "PostFlopGame":
        """isomorphism=False deals every card at every chance node (no suit
        isomorphism; the reference always applies it)."""
        lib = load_library()
        r0 = np.ascontiguousarray(card_config.range[0].data, np.float32)
        r1 = np.ascontiguousarray(card_config.range[1].data, np.float32)
        cc = _CardConfigC()
        cc.range[0] = _fptr(r0)
        cc.range[1] = _fptr(r1)
        for i in range(3):
            cc.flop[i] = card_config.flop[i]
        cc.turn = card_config.turn
        cc.river = card_config.river
        h = C.c_void_p()
        _check(lib.psr_game_build_ex(C.byref(cc), action_tree._h, 0 if isomorphism else 1, C.byref(h)))
        return cls(h)

    def __del__(self):
        if getattr(self, "_h", None):
            load_library().psr_game_free(self._h)
            self._h = None

    def num_private_hands(self, player: int) -> int:
        return _check(load_library().psr_game_num_hands(self._h, player))

    def private_cards(self, player: int) -> List[Tuple[int, int]]:
        n = self.num_private_hands(player)
        buf = (C.c_uint8 * (2 * n))()
        _check(load_library().psr_game_private_cards(self._h, player, buf))
        return [(buf[2 * i], buf[2 * i + 1]) for i in range(n)]

    def initial_weights(self, player: int) -> np.ndarray:
        out = np.zeros(self.num_private_hands(player), np.float32)
        _check(load_library().psr_game_initial_weights(self._h, player, _fptr(out)))
        return out

    def num_nodes(self) -> int:
        return _check(load_library().psr_game_num_nodes(self._h))

    def num_combinations(self) -> float:
        return load_library().psr_game_num_combinations(self._h)

    def num_storage(self) -> int:
        return _check(load_library().psr_game_num_storage(self._h))

    def num_boundary_nodes(self) -> int:
        """interpreter.rs:313-317: depth-boundary leaves of the tree."""
        return _check(load_library().psr_game_num_boundaries(self._h))

    def boundary_node_indices(self) -> list:
        """interpreter.rs:320-330: arena index of each boundary ordinal."""
        return [_check(load_library().psr_game_boundary_node(self._h, b)) for b in range(self.num_boundary_nodes())]

    def boundary_pot(self, ordinal: int) -> int:
        """interpreter.rs:354-358."""
        return _check(load_library().psr_game_boundary_pot(self._h, ordinal))

    def topology(self):
        """(parent, depth) of every node: arena indices are the node ids of
        gpu-range-solver's extract_topology (extract.rs:55-191) and edge e
        leads to node e + 1."""
        n = self.num_nodes()
        par = np.zeros(n, np.int32)
        dep = np.zeros(n, np.int32)
        _check(load_library().psr_game_topology(self._h, par.ctypes.data_as(C.POINTER(C.c_int32)),
                                                dep.ctypes.data_as(C.POINTER(C.c_int32)), n))
        return par, dep

    def validate_layout(self) -> dict:
        """Host-only check of the device layout (no GPU needed)."""
        out = (C.c_int64 * 6)()
        _check(load_library().psr_layout_validate(self._h, out))
        return dict(zip(["tiles", "levels", "terminals", "widest_level", "state", "nodes"], list(out)))

    def serialize(self) -> bytes:
        lib = load_library()
        n = _check(lib.psr_game_serialize(self._h, None, 0))
        buf = (C.c_uint8 * n)()
        _check(lib.psr_game_serialize(self._h, buf, n))
        return bytes(buf)


# ---------------------------------------------------------------- solving --

@dataclass
class GpuSolverConfig:
    """gpu-range-solver lib.rs:14-18 (+ check_interval, reference value 5)."""
    max_iterations: int = 1000
    target_exploitability: float = 0.0
    print_progress: bool = False
    check_interval: int = 5


@dataclass
class GpuSolveResult:
    """gpu-range-solver lib.rs:21-28."""
    exploitability: float
    iterations_run: int
    root_strategy: np.ndarray


def gpu_solve_hand_parallel(game: PostFlopGame, config: GpuSolverConfig) -> GpuSolveResult:
    """gpu-range-solver lib.rs:71-87: solve one game on the GPU."""
    lib = load_library()
    cfg = _SolverConfigC(config.max_iterations, config.target_exploitability, int(config.print_progress),
                         config.check_interval)
    res = _SolveResultC()
    cap = 64 * 1326
    strat = np.zeros(cap, np.float32)
    _check(lib.psr_gpu_solve(game._h, C.byref(cfg), C.byref(res), _fptr(strat), cap))
    return GpuSolveResult(res.exploitability, res.iterations_run, strat)


gpu_solve_game = gpu_solve_hand_parallel
gpu_solve_mega = gpu_solve_hand_parallel


STORAGE_MODES = {"f32": 0, "compressed": 1, "fp16": 2, "bf16": 3}


class GpuBatchSolver:
    """gpu-range-solver batch.rs:79-610, with one tree per game.

    storage: "f32" (default), "compressed" (the reference's i16/u16 with a
    per-node scale, allocate_memory(true)), "fp16" or "bf16" (value/scale
    words with a per-node f32 scale)."""

    def __init__(self, device: int = 0, storage: str = "f32"):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.psr_batch_create(device, C.byref(h)))
        self._h = h
        self.storage = storage
        _check(lib.psr_batch_set_storage(h, STORAGE_MODES[storage]))
        self.games: List[PostFlopGame] = []

    def __del__(self):
        if getattr(self, "_h", None):
            load_library().psr_batch_destroy(self._h)
            self._h = None

    def prepare_batch(self, games: Sequence[PostFlopGame], max_iterations: int,
                      targets: Optional[Sequence[float]] = None):
        lib = load_library()
        self.games = list(games)
        arr = (C.c_void_p * len(games))(*[g._h.value for g in games])
        t = None
        if targets is not None:
            tt = np.ascontiguousarray(targets, np.float32)
            self._targets = tt
            t = _fptr(tt)
        _check(lib.psr_batch_prepare(self._h, arr, len(games), max_iterations, t))

    def reset(self):
        _check(load_library().psr_batch_reset(self._h))

    def set_targets(self, targets: Optional[Sequence[float]]):
        """Early-stop targets (chips per game; None = 0) for the next solve."""
        t = None
        if targets is not None:
            self._targets = np.ascontiguousarray(targets, np.float32)
            t = _fptr(self._targets)
        _check(load_library().psr_batch_set_targets(self._h, t))

    def run_iterations(self, start: int, end: int):
        _check(load_library().psr_batch_run_iterations(self._h, start, end))

    def solve(self, max_iterations: int, check_interval: int = 5):
        n = len(self.games)
        expl = np.zeros(n, np.float32)
        iters = np.zeros(n, np.uint32)
        _check(load_library().psr_batch_solve(self._h, max_iterations, check_interval, _fptr(expl),
                                              iters.ctypes.data_as(C.POINTER(C.c_uint32))))
        return expl, iters

    def exploitability(self) -> np.ndarray:
        expl = np.zeros(len(self.games), np.float32)
        _check(load_library().psr_batch_exploitability(self._h, _fptr(expl)))
        return expl

    def exploitability_f64(self) -> np.ndarray:
        """FP64 best-response exploitability of every game (device)."""
        out = np.zeros(len(self.games), np.float64)
        _check(load_library().psr_batch_exploitability_f64(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def compute_root_cfvalues(self):
        """utility.rs:348-372 finalize: the cfvalue pass under the average
        strategy (also records the reach at every depth-boundary leaf)."""
        _check(load_library().psr_batch_compute_root_cfvalues(self._h))

    def root_cfvalues(self, compute: bool = True) -> List[Tuple[np.ndarray, np.ndarray]]:
        """Per-hand root cfvalues of every game; compute=False reads the ones
        the last solve() (which ends with finalize) or cfvalue pass left."""
        lib = load_library()
        if compute:
            _check(lib.psr_batch_compute_root_cfvalues(self._h))
        out = []
        for i, g in enumerate(self.games):
            pair = []
            for p in range(2):
                v = np.zeros(g.num_private_hands(p), np.float32)
                _check(lib.psr_batch_root_cfvalues(self._h, i, p, _fptr(v)))
                pair.append(v)
            out.append(tuple(pair))
        return out

    def root_cfvalues_all(self, player: int) -> np.ndarray:
        """Root cfvalues of `player` of every game, concatenated (as the last
        solve() or cfvalue pass left them)."""
        n = sum(g.num_private_hands(player) for g in self.games)
        out = np.zeros(n, np.float32)
        _check(load_library().psr_batch_root_cfvalues_all(self._h, player, _fptr(out), n))
        return out

    def set_initial_weights(self, w0, w1, game: int = -1):
        """New initial weights (root reach) of one game or of every game
        (game = -1: concatenated in batch order) on the same trees."""
        a = np.ascontiguousarray(w0, np.float32)
        b = np.ascontiguousarray(w1, np.float32)
        _check(load_library().psr_batch_set_initial_weights(self._h, game, _fptr(a), len(a), _fptr(b), len(b)))

    def strategy_sum(self, i: int) -> np.ndarray:
        n = self._game(i).num_storage()
        v = np.zeros(n, np.float32)
        _check(load_library().psr_batch_strategy_sum(self._h, i, _fptr(v), n))
        return v

    def regrets(self, i: int) -> np.ndarray:
        n = self._game(i).num_storage()
        v = np.zeros(n, np.float32)
        _check(load_library().psr_batch_regrets(self._h, i, _fptr(v), n))
        return v

    def raw_state(self, i: int):
        """16-bit words and per-node scales: (regret u16, strategy u16,
        regret scales, strategy scales)."""
        n = self._game(i).num_storage()
        r = np.zeros(n, np.uint16)
        s = np.zeros(n, np.uint16)
        cap = 1 << 20
        rs = np.zeros(cap, np.float32)
        ss = np.zeros(cap, np.float32)
        k = _check(load_library().psr_batch_raw_state(self._h, i, r.ctypes.data_as(C.POINTER(C.c_uint16)),
                                                      s.ctypes.data_as(C.POINTER(C.c_uint16)), _fptr(rs), _fptr(ss),
                                                      n, cap))
        return r, s, rs[:k], ss[:k]

    def normalized_weights(self, i: int, player: int) -> np.ndarray:
        """query.rs:498-567 at the root of game i."""
        out = np.zeros(self._game(i).num_private_hands(player), np.float32)
        _check(load_library().psr_batch_normalized_weights(self._h, i, player, _fptr(out), len(out)))
        return out

    def expected_values(self, i: int, player: int) -> np.ndarray:
        """query.rs:639-666 at the root of game i (chips per hand)."""
        out = np.zeros(self._game(i).num_private_hands(player), np.float32)
        _check(load_library().psr_batch_expected_values(self._h, i, player, _fptr(out), len(out)))
        return out

    def expected_values_detail(self, i: int, player: int) -> np.ndarray:
        """query.rs:672-760 at the root of game i (per action and hand for the
        player acting at the root)."""
        out = np.zeros(16 * self._game(i).num_private_hands(player), np.float32)
        n = _check(load_library().psr_batch_expected_values_detail(self._h, i, player, _fptr(out), len(out)))
        return out[:n].copy()

    def _game(self, i: int):
        if not 0 <= i < len(self.games):
            raise PsrError(f"game index {i} out of range")
        return self.games[i]

    def num_boundaries(self, i: int) -> int:
        return _check(load_library().psr_batch_num_boundaries(self._h, i))

    def set_boundary_cfvs(self, i: int, ordinal: int, player: int, cfvs, raw: bool = False):
        """interpreter.rs:343-352 set_boundary_cfvs on game i of the batch;
        raw=True: chip cfvalues used as they are (evaluation.rs:606-630)."""
        v = np.ascontiguousarray(cfvs, np.float32)
        _check(load_library().psr_batch_set_boundary_cfvs(self._h, i, ordinal, player, _fptr(v), len(v),
                                                          1 if raw else 0))

    def update_leaf_cfvs(self, p0, p1):
        """batch.rs:557-585: raw leaf values, per player: after
        set_leaf_injection [game][injected leaf][num_hands] as the reference
        lays them out; otherwise every boundary of every game (per game in
        batch order, per ordinal, H_player values)."""
        a = np.ascontiguousarray(p0, np.float32)
        b = np.ascontiguousarray(p1, np.float32)
        _check(load_library().psr_batch_update_leaf_cfvs(self._h, _fptr(a), len(a), _fptr(b), len(b)))

    def set_leaf_injection(self, node_ids, depths, num_hands: int):
        """batch.rs:279-300: leaves whose values update_leaf_cfvs then takes
        as [game][leaf][num_hands] arrays (num_hands 1326 = combo space)."""
        ids = np.ascontiguousarray(node_ids, np.int32)
        dep = np.ascontiguousarray(depths, np.int32)
        _check(load_library().psr_batch_set_leaf_injection(self._h, ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                                           dep.ctypes.data_as(C.POINTER(C.c_int32)), len(ids),
                                                           num_hands))

    def download_reach(self, num_hands: int) -> np.ndarray:
        """batch.rs:537-546: [sum of nodes over games, num_hands] reach of
        player 0 from the last traversal."""
        n = sum(g.num_nodes() for g in self.games)
        out = np.zeros(n * num_hands, np.float32)
        _check(load_library().psr_batch_download_reach(self._h, _fptr(out), len(out), num_hands))
        return out.reshape(n, num_hands)

    def extract_results(self, num_hands: int) -> List[np.ndarray]:
        """batch.rs:587-609: per game the strategy sums as [edges, num_hands]."""
        lib = load_library()
        res = []
        for i, g in enumerate(self.games):
            out = np.zeros((g.num_nodes() - 1) * num_hands, np.float32)
            _check(lib.psr_batch_strategy_sum_edges(self._h, i, _fptr(out), len(out), num_hands))
            res.append(out.reshape(-1, num_hands))
        return res

    def boundary_reach(self, i: int, ordinal: int, player: int) -> np.ndarray:
        """interpreter.rs:396-403: reach of `player` recorded at the last
        visit of the boundary by the opponent."""
        out = np.zeros(self._game(i).num_private_hands(player), np.float32)
        _check(load_library().psr_batch_boundary_reach(self._h, i, ordinal, player, _fptr(out), len(out)))
        return out

    def reach_at_nodes(self, i: int, player: int, nodes) -> np.ndarray:
        """batch.rs:765-797 compute_reach_at_nodes: reach of `player` at the
        given arena nodes under the average strategy, (len(nodes), H_player).
        The reference's result[p] is reach_at_nodes(i, 1 - p, nodes)."""
        g = self._game(i)
        nodes = np.ascontiguousarray(nodes, np.int64)
        out = np.zeros((len(nodes), g.num_private_hands(player)), np.float32)
        _check(load_library().psr_batch_reach_at_nodes(self._h, i, player,
                                                       nodes.ctypes.data_as(C.POINTER(C.c_int64)), len(nodes),
                                                       _fptr(out)))
        return out

    def last_kernel_ms(self) -> float:
        return float(load_library().psr_batch_last_kernel_ms(self._h))

    def set_launch(self, graph_iterations: int, subtree_per_cu: int, blocks_per_cu: int):
        """Launch shape: DCFR iterations per replayed graph of the multi-tile
        games (0 = eager launches), subtree workgroups per CU, whole-game
        workgroups per CU."""
        _check(load_library().psr_batch_set_launch(self._h, graph_iterations, subtree_per_cu, blocks_per_cu))

    def autotune(self, iterations: int = 10) -> dict:
        """Times the candidate launch shapes over `iterations` DCFR
        iterations, applies the fastest (cached per workload) and resets the
        solve; returns the shape and whether it came from the cache."""
        out = (C.c_int32 * 3)()
        rc = _check(load_library().psr_batch_autotune(self._h, iterations, out))
        return {"graph_iterations": out[0], "subtree_per_cu": out[1], "blocks_per_cu": out[2], "cached": rc == 1}

    def launch_info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(load_library().psr_batch_launch_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"num_cus": a.value, "grid": b.value, "blocks_per_cu": c.value}

    def set_depth_first(self, enable: bool):
        """Depth-first traversals with the rows in LDS for the single-tile
        games that fit (default), or the level-by-level path for every game;
        results are identical."""
        _check(load_library().psr_batch_set_depth_first(self._h, 1 if enable else 0))

    def depth_first_info(self) -> dict:
        """Tiles on the depth-first path (single-tile games and the river
        subtree tiles of turn and flop games), those of them that run one per
        wave (subtree tiles of games with at most 256 hands per player), and
        the LDS of the whole-workgroup launches."""
        n, nw, lds = C.c_int32(), C.c_int32(), C.c_int64()
        _check(load_library().psr_batch_depth_first_info(self._h, C.byref(n), C.byref(nw), C.byref(lds)))
        return {"tiles": n.value, "wave_tiles": nw.value, "lds_bytes": lds.value}
