"""Training-record format of the reference's datagen (cfvnet/src/datagen/
storage.rs), mirroring its tests (storage.rs:126-260): round trips of one and
several records, record sizes for 4- and 5-card boards, record counts; and
the byte layout of the reference test's sample record checked at fixed
offsets computed from the documented field order."""
import io
import struct

import numpy as np

from poker_solver_rust_amd import records as RC


def sample_record():
    """storage.rs:126-149 sample_record."""
    oop = np.zeros(1326, np.float32)
    ip = np.zeros(1326, np.float32)
    cfvs = np.zeros(1326, np.float32)
    mask = np.zeros(1326, np.uint8)
    oop[0] = 0.5
    oop[1] = 0.5
    ip[100] = 1.0
    cfvs[0] = 0.123
    cfvs[100] = -0.456
    mask[[0, 1, 100]] = 1
    return RC.TrainingRecord([0, 4, 8, 12, 16], 100.0, 50.0, 0, 0.05, oop, ip, cfvs, mask)


def test_record_size_is_consistent():
    assert RC.record_size(5) == 1 + 5 + 4 + 4 + 1 + 4 + 1326 * 4 * 3 + 1326 == 17257
    assert RC.record_size(4) == 17256


def test_byte_layout_of_sample_record():
    f = io.BytesIO()
    RC.write_record(f, sample_record())
    b = f.getvalue()
    assert len(b) == RC.record_size(5)
    assert b[0] == 5 and list(b[1:6]) == [0, 4, 8, 12, 16]
    assert struct.unpack_from("<f", b, 6)[0] == 100.0
    assert struct.unpack_from("<f", b, 10)[0] == 50.0
    assert b[14] == 0
    assert struct.unpack_from("<f", b, 15)[0] == np.float32(0.05)
    base = 19
    assert struct.unpack_from("<f", b, base + 4 * 1)[0] == 0.5                     # oop_range[1]
    assert struct.unpack_from("<f", b, base + 1326 * 4 + 4 * 100)[0] == 1.0        # ip_range[100]
    assert struct.unpack_from("<f", b, base + 2 * 1326 * 4 + 4 * 100)[0] == np.float32(-0.456)   # cfvs[100]
    m = base + 3 * 1326 * 4
    assert b[m] == 1 and b[m + 1] == 1 and b[m + 2] == 0 and b[m + 100] == 1


def test_round_trip_single_and_multiple():
    r1 = sample_record()
    r2 = sample_record()
    r2.player = 1
    r2.pot = 200.0
    r2.game_value = -0.1
    f = io.BytesIO()
    RC.write_record(f, r1)
    RC.write_record(f, r2)
    f.seek(0)
    a, b = RC.read_records(f)
    for x, y in ((r1, a), (r2, b)):
        assert x.board == y.board and x.pot == y.pot and x.effective_stack == y.effective_stack
        assert x.player == y.player and abs(x.game_value - y.game_value) < 1e-7
        for u, v in ((x.oop_range, y.oop_range), (x.ip_range, y.ip_range), (x.cfvs, y.cfvs)):
            assert np.array_equal(u, v)
        assert np.array_equal(x.valid_mask, y.valid_mask)
    assert RC.count_records(len(f.getvalue()), 5) == 2


def test_four_card_board_and_truncation():
    r = sample_record()
    r.board = [0, 4, 8, 12]
    f = io.BytesIO()
    RC.write_record(f, r)
    data = f.getvalue()
    assert len(data) == RC.record_size(4)
    assert RC.read_record(io.BytesIO(data)).board == [0, 4, 8, 12]
    try:
        RC.read_record(io.BytesIO(data[:-3]))
        raise AssertionError("truncated record read")
    except EOFError:
        pass
