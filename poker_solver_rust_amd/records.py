"""Training records of solved subgames in the reference's binary format
(cfvnet/src/datagen/storage.rs:1-125, read by cfvnet/python/cfvnet/data.py),
so that river solves from the batched device solver feed the reference's
value-net training unchanged.

Record layout, little-endian, one per (subgame, player):
    board_size u8, board u8[board_size], pot f32, effective_stack f32,
    player u8, game_value f32, oop_range f32[1326], ip_range f32[1326],
    cfvs f32[1326], valid_mask u8[1326]

Values (domain/solver.rs:33-95, domain/pipeline.rs:356-413): for every hand of
player p, cfvs[combo] = (EV_chips - pot / 2) / (pot / 2) with EV the query
interface's expected value (query.rs:639-666); valid_mask marks the combos
of either player's hand list; game_value = sum over combos of
range_p * cfvs in f32, summed in combo order.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass
from typing import BinaryIO, List, Sequence

import numpy as np

NUM_COMBOS = 1326


def record_size(board_size: int) -> int:
    """storage.rs:10-21 record_size."""
    return 1 + board_size + 4 + 4 + 1 + 4 + NUM_COMBOS * 4 * 3 + NUM_COMBOS


@dataclass
class TrainingRecord:
    """storage.rs:24-35 TrainingRecord."""
    board: List[int]
    pot: float
    effective_stack: float
    player: int
    game_value: float
    oop_range: np.ndarray     # f32[1326]
    ip_range: np.ndarray      # f32[1326]
    cfvs: np.ndarray          # f32[1326]
    valid_mask: np.ndarray    # u8[1326]


def write_record(f: BinaryIO, rec: TrainingRecord) -> None:
    """storage.rs:43-56 write_record."""
    board = bytes(int(c) for c in rec.board)
    f.write(struct.pack("<B", len(board)))
    f.write(board)
    f.write(struct.pack("<ffBf", rec.pot, rec.effective_stack, rec.player, rec.game_value))
    for a in (rec.oop_range, rec.ip_range, rec.cfvs):
        arr = np.ascontiguousarray(a, "<f4")
        if arr.shape != (NUM_COMBOS,):
            raise ValueError("range / cfv arrays hold 1326 values")
        f.write(arr.tobytes())
    mask = np.ascontiguousarray(rec.valid_mask, np.uint8)
    if mask.shape != (NUM_COMBOS,):
        raise ValueError("valid_mask holds 1326 values")
    f.write(mask.tobytes())


def read_record(f: BinaryIO) -> TrainingRecord:
    """storage.rs:61-107 read_record; raises EOFError at the end of the file."""
    head = f.read(1)
    if len(head) == 0:
        raise EOFError
    n = head[0]
    body = f.read(record_size(n) - 1)
    if len(body) != record_size(n) - 1:
        raise EOFError("truncated record")
    board = list(body[:n])
    pot, stack, player, gv = struct.unpack_from("<ffBf", body, n)
    off = n + 13
    arrs = []
    for _ in range(3):
        arrs.append(np.frombuffer(body, "<f4", NUM_COMBOS, off).astype(np.float32))
        off += NUM_COMBOS * 4
    mask = np.frombuffer(body, np.uint8, NUM_COMBOS, off).copy()
    return TrainingRecord(board, pot, stack, player, gv, arrs[0], arrs[1], arrs[2], mask)


def read_records(f: BinaryIO) -> List[TrainingRecord]:
    out = []
    while True:
        try:
            out.append(read_record(f))
        except EOFError:
            return out


def count_records(nbytes: int, board_size: int) -> int:
    """storage.rs:112-116 count_records (uniform board size)."""
    return nbytes // record_size(board_size)


def _f32_sum_in_order(x: np.ndarray) -> np.float32:
    """Sequential f32 sum (Rust's Iterator::sum over f32), combo order."""
    return np.cumsum(np.asarray(x, np.float32), dtype=np.float32)[-1]


def records_from_batch(solver, subgames: Sequence, first: int = 0) -> List[TrainingRecord]:
    """Two records (OOP, IP) per solved subgame of a GpuBatchSolver, as the
    reference's datagen writes them (domain/solver.rs:33-95).  `subgames`
    are the workloads.Subgame descriptions of games first, first + 1, ... of
    the batch (board, 1326-combo ranges, pot, stack)."""
    out = []
    for k, s in enumerate(subgames):
        i = first + k
        g = solver.games[i]
        half = 0.5 * float(s.pot)
        norm = half if half > 0.0 else 1.0
        cf = [np.zeros(NUM_COMBOS, np.float32) for _ in range(2)]
        mask = np.zeros(NUM_COMBOS, np.uint8)
        for p in range(2):
            ev = solver.expected_values(i, p)
            pc = np.asarray(g.private_cards(p), np.int64).reshape(-1, 2)
            c1, c2 = pc.min(1), pc.max(1)
            idx = c1 * (101 - c1) // 2 + c2 - 1          # card_pair_to_index, card.rs:50-57
            cf[p][idx] = ((ev.astype(np.float64) - half) / norm).astype(np.float32)
            mask[idx] = 1
        ranges = [np.asarray(s.ranges[p], np.float32) for p in range(2)]
        board = [int(c) for c in s.board]
        for p in range(2):
            gv = _f32_sum_in_order(ranges[p] * cf[p])
            out.append(TrainingRecord(board, float(s.pot), float(s.stack), p, float(gv), ranges[0], ranges[1],
                                      cf[p], mask))
    return out
