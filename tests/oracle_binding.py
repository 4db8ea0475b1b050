"""ctypes binding of oracle/liboracle.so (the CPU oracle).  Test
infrastructure: used by tests/, __graft_entry__.smoke() as the checker and
bench.py's cpu_baseline leg only."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
_LIB = None


def build_oracle():
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(ORACLE_LIB):
        build_oracle()
    lib = C.CDLL(ORACLE_LIB)
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    sig = {
        "oracle_last_error": (C.c_char_p, []),
        "oracle_game_load": (vp, [C.POINTER(C.c_uint8), C.c_size_t, C.c_int]),
        "oracle_game_free": (None, [vp]),
        "oracle_storage_size": (C.c_int64, [vp]),
        "oracle_num_hands": (C.c_int, [vp, C.c_int]),
        "oracle_discount": (None, [C.c_uint32, C.POINTER(C.c_float)]),
        "oracle_solve_step": (None, [vp, C.c_uint32]),
        "oracle_exploitability": (C.c_double, [vp]),
        "oracle_solve": (C.c_double, [vp, C.c_uint32, C.c_double, C.POINTER(C.c_uint32)]),
        "oracle_mes_ev": (None, [vp, dp]),
        "oracle_current_ev": (None, [vp, dp]),
        "oracle_root_cfvalues": (None, [vp, C.c_int, dp]),
        "oracle_state": (None, [vp, dp, dp]),
        "oracle_set_state": (None, [vp, dp, dp]),
        "oracle_evaluate": (C.c_int, [vp, C.c_uint32, C.c_int, dp, dp]),
        "oracle_reach_at_node": (C.c_int, [vp, C.c_uint32, C.c_int, dp]),
        "oracle_set_boundary_cfvs": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int, C.c_int]),
        "oracle_boundary_reach": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int]),
        "oracle_num_boundaries": (C.c_int, [vp]),
        "oracle_expected_values": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_float), C.c_int]),
        "oracle_normalized_weights": (None, [vp, C.c_int, C.POINTER(C.c_float)]),
        "oracle_compressed_state": (C.c_int, [vp, C.POINTER(C.c_int16), C.POINTER(C.c_uint16),
                                              C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "oracle_num_decisions": (C.c_int, [vp]),
        "oracle_encode_signed": (C.c_float, [C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_int16)]),
        "oracle_encode_unsigned": (C.c_float, [C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_uint16)]),
        "oracle_regret_matching": (None, [C.POINTER(C.c_float), C.c_size_t, C.c_int, C.POINTER(C.c_float)]),
        "oracle_regret_matching_compressed": (None, [C.POINTER(C.c_int16), C.c_size_t, C.c_int,
                                                     C.POINTER(C.c_float)]),
        "oracle_slice_op": (C.c_int, [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                      C.c_size_t, C.c_size_t, C.c_float]),
        "oracle_debug_break_isomorphism": (C.c_int, [vp, C.c_int]),
        "oracle_set_threads": (None, [vp, C.c_int]),
        "oracle_node_player": (C.c_int, [vp, C.c_uint32]),
        "oracle_solve_many": (C.c_double, [C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t), C.c_int,
                                           C.c_uint32, C.c_int, dp, dp, C.POINTER(C.c_uint32)]),
        "oracle_fp64_expl_many": (C.c_double, [C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t), C.c_int,
                                               C.POINTER(C.c_uint32), C.c_int, dp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def discount(t):
    out = (C.c_float * 3)()
    load().oracle_discount(t, out)
    return tuple(out)


class OracleGame:
    """One game loaded from PostFlopGame.serialize(); precision 32 follows the
    reference's f32 arithmetic, 64 is the all-f64 oracle."""

    def __init__(self, blob: bytes, precision: int = 32):
        lib = load()
        self._blob = (C.c_uint8 * len(blob)).from_buffer_copy(blob)
        self._h = lib.oracle_game_load(self._blob, len(blob), precision)
        if not self._h:
            raise RuntimeError(lib.oracle_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            load().oracle_game_free(self._h)
            self._h = None

    def num_hands(self, p):
        return load().oracle_num_hands(self._h, p)

    def storage_size(self):
        return load().oracle_storage_size(self._h)

    def solve_step(self, t):
        load().oracle_solve_step(self._h, t)

    def set_threads(self, threads):
        """threads > 1: the children of nodes before the river deal run in
        parallel (the reference's rayon policy); results are unchanged."""
        load().oracle_set_threads(self._h, threads)

    def exploitability(self):
        return load().oracle_exploitability(self._h)

    def solve(self, max_iterations, target=0.0):
        it = C.c_uint32()
        e = load().oracle_solve(self._h, max_iterations, target, C.byref(it))
        return e, it.value

    def mes_ev(self):
        out = np.zeros(2)
        load().oracle_mes_ev(self._h, _dptr(out))
        return out

    def current_ev(self):
        out = np.zeros(2)
        load().oracle_current_ev(self._h, _dptr(out))
        return out

    def root_cfvalues(self, p):
        out = np.zeros(self.num_hands(p))
        load().oracle_root_cfvalues(self._h, p, _dptr(out))
        return out

    def state(self):
        n = self.storage_size()
        r, s = np.zeros(n), np.zeros(n)
        load().oracle_state(self._h, _dptr(r), _dptr(s))
        return r, s

    def set_state(self, regret, strategy):
        r = np.ascontiguousarray(regret, np.float64)
        s = np.ascontiguousarray(strategy, np.float64)
        load().oracle_set_state(self._h, _dptr(r), _dptr(s))

    def reach_at_node(self, node, player):
        out = np.zeros(self.num_hands(player))
        rc = load().oracle_reach_at_node(self._h, node, player, _dptr(out))
        assert rc == 0, "node not reached"
        return out

    def evaluate(self, node, player, cfreach):
        cr = np.ascontiguousarray(cfreach, np.float64)
        out = np.zeros(self.num_hands(player))
        rc = load().oracle_evaluate(self._h, node, player, _dptr(cr), _dptr(out))
        if rc != 0:
            raise ValueError("not a terminal node")
        return out

    def compressed_state(self):
        """precision-16 handle: (i16 regrets, u16 strategy, regret scales,
        strategy scales), decision nodes in arena order."""
        lib = load()
        n = self.storage_size()
        k = lib.oracle_num_decisions(self._h)
        r = np.zeros(n, np.int16)
        s = np.zeros(n, np.uint16)
        rs = np.zeros(k, np.float32)
        ss = np.zeros(k, np.float32)
        rc = lib.oracle_compressed_state(self._h, r.ctypes.data_as(C.POINTER(C.c_int16)),
                                         s.ctypes.data_as(C.POINTER(C.c_uint16)),
                                         rs.ctypes.data_as(C.POINTER(C.c_float)), ss.ctypes.data_as(C.POINTER(C.c_float)))
        if rc != 0:
            raise ValueError("not a compressed handle")
        return r, s, rs, ss

    def set_boundary_cfvs(self, ordinal, player, values, raw=False):
        v = np.ascontiguousarray(values, np.float32)
        rc = load().oracle_set_boundary_cfvs(self._h, ordinal, player, v.ctypes.data_as(C.POINTER(C.c_float)), len(v),
                                             1 if raw else 0)
        if rc != 0:
            raise ValueError("bad boundary ordinal or length")

    def expected_values(self, player, detail=False):
        """query.rs:639-760 at the root (chips per hand; per action and hand
        with detail=True for the player acting at the root)."""
        out = np.zeros(16 * 1326, np.float32)
        n = load().oracle_expected_values(self._h, player, 1 if detail else 0, out.ctypes.data_as(C.POINTER(C.c_float)),
                                          len(out))
        if n < 0:
            raise ValueError("no decision node at the root")
        return out[:n].copy()

    def normalized_weights(self, player):
        out = np.zeros(self.num_hands(player), np.float32)
        load().oracle_normalized_weights(self._h, player, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def num_boundaries(self):
        return load().oracle_num_boundaries(self._h)

    def node_kinds(self):
        """Per node (arena order): (kind, player) with kind one of
        "decision", "chance", "fold", "showdown", "boundary"."""
        out = []
        n = 0
        while True:
            f = load().oracle_node_player(self._h, n)
            if f < 0:
                return out
            if f & 8:
                kind = "boundary" if (f & 40) == 40 else "fold" if (f & 24) == 24 else "showdown"
            else:
                kind = "chance" if f & 4 else "decision"
            out.append((kind, f & 3))
            n += 1

    def break_isomorphism(self, mode):
        """Plants a mistake in the isomorphic chance sums (1: no hand swaps,
        2: isomorphic cards left out); returns the entries changed."""
        return load().oracle_debug_break_isomorphism(self._h, mode)

    def boundary_reach(self, ordinal, player):
        """Opponent reach recorded at the last visit of boundary `ordinal`,
        indexed like the reference's boundary_reach(ordinal, player)."""
        out = np.zeros(1326, np.float32)
        n = load().oracle_boundary_reach(self._h, ordinal, player, out.ctypes.data_as(C.POINTER(C.c_float)), 1326)
        if n < 0:
            raise ValueError("bad boundary ordinal")
        return out[:n].copy()


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def regret_matching(regret, num_actions):
    """solver.rs:657-662 regret_matching (f32)."""
    r = np.ascontiguousarray(regret, np.float32)
    out = np.zeros(len(r), np.float32)
    load().oracle_regret_matching(_fp(r), len(r), num_actions, _fp(out))
    return out


def regret_matching_compressed(regret, num_actions):
    """solver.rs:667-677 regret_matching_compressed (i16 regrets)."""
    r = np.ascontiguousarray(regret, np.int16)
    out = np.zeros(len(r), np.float32)
    load().oracle_regret_matching_compressed(r.ctypes.data_as(C.POINTER(C.c_int16)), len(r), num_actions, _fp(out))
    return out


_SLICE_OPS = {"sub": 0, "mul": 1, "div": 2, "sum": 3, "fma": 4, "max_slices": 5, "max": 6}


def slice_op(op, lhs, rhs=None, src2=None, row=None, default=0.0):
    """One of the oracle's slice operations (sliceop.rs); returns the result.
    sub / mul / div act element-wise on lhs; sum / fma / max_slices reduce
    the rows (of `row` values) of rhs (and src2) into a new row; max is the
    scalar max(lhs, rhs)."""
    code = _SLICE_OPS[op]
    if op in ("sum", "fma", "max_slices"):
        src = np.ascontiguousarray(rhs, np.float32)
        s2 = np.ascontiguousarray(src if src2 is None else src2, np.float32)
        dst = np.zeros(row, np.float32)
        rc = load().oracle_slice_op(code, _fp(dst), _fp(src), _fp(s2), row, len(src), 0.0)
        assert rc == 0
        return dst
    if op == "max":
        a = np.array([lhs], np.float32)
        b = np.array([rhs], np.float32)
        load().oracle_slice_op(code, _fp(a), _fp(b), _fp(b), 1, 1, 0.0)
        return float(a[0])
    a = np.array(lhs, np.float32)
    b = np.ascontiguousarray(rhs, np.float32)
    load().oracle_slice_op(code, _fp(a), _fp(b), _fp(b), len(a), len(a), default)
    return a


def encode_signed(values):
    v = np.ascontiguousarray(values, np.float32)
    out = np.zeros(len(v), np.int16)
    sc = load().oracle_encode_signed(v.ctypes.data_as(C.POINTER(C.c_float)), len(v),
                                     out.ctypes.data_as(C.POINTER(C.c_int16)))
    return sc, out


def encode_unsigned(values):
    v = np.ascontiguousarray(values, np.float32)
    out = np.zeros(len(v), np.uint16)
    sc = load().oracle_encode_unsigned(v.ctypes.data_as(C.POINTER(C.c_float)), len(v),
                                       out.ctypes.data_as(C.POINTER(C.c_uint16)))
    return sc, out


def _blob_array(blobs):
    n = len(blobs)
    bufs = [(C.c_uint8 * len(b)).from_buffer_copy(b) for b in blobs]
    ptrs = (C.POINTER(C.c_uint8) * n)(*[C.cast(b, C.POINTER(C.c_uint8)) for b in bufs])
    lens = (C.c_size_t * n)(*[len(b) for b in blobs])
    return bufs, ptrs, lens


def solve_many(blobs, iterations, threads, targets=None):
    """Solves serialized games with the reference solve() schedule (which
    ends with finalize) using `threads` worker threads (one game per thread
    at a time).  Returns (wall seconds, exploitability array, iterations
    array)."""
    lib = load()
    n = len(blobs)
    bufs, ptrs, lens = _blob_array(blobs)
    expl = np.zeros(n)
    its = np.zeros(n, np.uint32)
    tg = None if targets is None else np.ascontiguousarray(targets, np.float64)
    secs = lib.oracle_solve_many(ptrs, lens, n, iterations, threads, None if tg is None else _dptr(tg), _dptr(expl),
                                 its.ctypes.data_as(C.POINTER(C.c_uint32)))
    del bufs
    return secs, expl, its


def fp64_exploitability_many(blobs, iterations, threads):
    """The FP64 oracle's exploitability of each game after exactly
    iterations[i] DCFR iterations (no early stop), `threads` games at a time.
    Returns (wall seconds, exploitability array)."""
    lib = load()
    n = len(blobs)
    bufs, ptrs, lens = _blob_array(blobs)
    its = np.ascontiguousarray(iterations, np.uint32)
    expl = np.zeros(n)
    secs = lib.oracle_fp64_expl_many(ptrs, lens, n, its.ctypes.data_as(C.POINTER(C.c_uint32)), threads, _dptr(expl))
    del bufs
    return secs, expl
