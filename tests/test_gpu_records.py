"""Training records written from a batched device solve (the reference
datagen's record contents, cfvnet/src/datagen/domain/solver.rs:33-95): the
values are the device expected values (checked against the oracle in
test_gpu_parity) turned into pot-relative cfvs, and the file reads back."""
import io

import numpy as np
import pytest

import oracle_binding as O
import poker_solver_rust_amd as P
from poker_solver_rust_amd import records as RC
from poker_solver_rust_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_records_from_batch():
    subs = W.river_batch(6, seed=31, iterations=100)
    games = [s.build() for s in subs]
    b = P.GpuBatchSolver(0)
    b.prepare_batch(games, 100, [s.target for s in subs])
    b.solve(100, 5)
    recs = RC.records_from_batch(b, subs)
    assert len(recs) == 2 * len(subs)
    f = io.BytesIO()
    for r in recs:
        RC.write_record(f, r)
    assert len(f.getvalue()) == len(recs) * RC.record_size(5)
    f.seek(0)
    back = RC.read_records(f)
    for k, (s, g) in enumerate(zip(subs, games)):
        half = 0.5 * s.pot
        for p in range(2):
            r = recs[2 * k + p]
            assert r.player == p and r.board == list(s.board) and r.pot == s.pot
            ev = b.expected_values(k, p)
            pc = np.asarray(g.private_cards(p)).reshape(-1, 2)
            idx = np.array([P.card_pair_to_index(int(a), int(c)) for a, c in pc])
            assert np.array_equal(r.cfvs[idx], ((ev.astype(np.float64) - half) / half).astype(np.float32))
            assert r.valid_mask[idx].all()
            gv = np.float32(0.0)
            for x in (np.asarray(s.ranges[p], np.float32) * r.cfvs):
                gv = np.float32(gv + x)
            assert np.float32(r.game_value) == gv
            assert np.array_equal(back[2 * k + p].cfvs, r.cfvs)
            assert back[2 * k + p].game_value == np.float32(r.game_value)
    # the OOP values of the first subgame against the oracle's expected values
    o = O.OracleGame(games[0].serialize())
    o.set_state(b.regrets(0).astype(np.float64), b.strategy_sum(0).astype(np.float64))
    ev0 = np.asarray(o.expected_values(0), np.float64)
    assert np.allclose(b.expected_values(0, 0), ev0, rtol=1e-4, atol=1e-3)


def test_river_datagen_end_to_end():
    """Sampled situations (R(S, p) ranges) -> batched device solve -> records
    (datagen/domain/solver.rs:33-95 contents): finite values on the valid
    combos only, game values finite, every situation converged or capped."""
    from poker_solver_rust_amd import datagen as D

    cfg = D.DatagenConfig(solver_iterations=200)
    recs, sits = D.generate_river_records(24, cfg, seed=5, chunk=16)
    assert len(recs) == 48 and len(sits) == 24
    for k, r in enumerate(recs):
        s = sits[k // 2]
        assert r.board == s.board and r.player == k % 2
        assert np.all(np.isfinite(r.cfvs)) and np.isfinite(r.game_value)
        assert np.all(r.cfvs[r.valid_mask == 0] == 0)
        # pot-relative values of a river subgame lie within [-1 - 2 stack / pot, 1 + 2 stack / pot]
        bound = 1.0 + 2.0 * s.stack / s.pot + 1e-3
        assert np.all(np.abs(r.cfvs) <= bound)
