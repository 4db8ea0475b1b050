"""psr_gpu_solve_noalloc (the allocation-free one-game solve) through the C
ABI, against psr_gpu_solve on the same games: two calls in a row, the second
with a smaller game (the scratch array keeps the first call's contents, so
the second shows that every buffer it reads is re-initialised), one on a
stream of the caller's, and a game too large for the array rejected with a
clear error."""
import ctypes as C

import numpy as np
import pytest

import poker_solver_rust_amd as P
from poker_solver_rust_amd import workloads as W
from poker_solver_rust_amd.range_solver import _SolverConfigC, _SolveResultC, _fptr, load_library

pytestmark = pytest.mark.gpu

CAP = 16 * 1326


def _lib():
    lib = load_library()
    fn = lib.psr_gpu_solve_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_void_p, C.POINTER(_SolverConfigC), C.POINTER(_SolveResultC), C.POINTER(C.c_float), C.c_int64,
                   C.c_void_p]
    return lib


def _solve(game, iters, target, noalloc, stream=None):
    lib = _lib()
    cfg = _SolverConfigC(iters, target, 0, 5)
    res = _SolveResultC()
    strat = np.full(CAP, -1.0, np.float32)
    if noalloc:
        rc = lib.psr_gpu_solve_noalloc(game._h, C.byref(cfg), C.byref(res), _fptr(strat), CAP, stream)
    else:
        rc = lib.psr_gpu_solve(game._h, C.byref(cfg), C.byref(res), _fptr(strat), CAP)
    if rc < 0:
        raise P.PsrError(lib.psr_last_error().decode())
    return res.exploitability, res.iterations_run, strat


def _std_game():
    cc = P.CardConfig(range=(P.Range.parse("AA,KK,QQ,AKs"), P.Range.parse("QQ-JJ,AQs,AJs")),
                      flop=P.flop_from_str("Qs Jh 2c"), turn=P.card_from_str("8d"), river=P.card_from_str("3s"))
    tc = P.TreeConfig(initial_state=P.BoardState.River, starting_pot=100, effective_stack=100,
                      river_bet_sizes=[("50%, a", ""), ("50%, a", "")])
    return P.PostFlopGame.with_config(cc, P.ActionTree(tc))


def _largest_batch_game():
    subs = W.river_batch(64, seed=7)
    games = [s.build() for s in subs]
    k = int(np.argmax([g.num_storage() for g in games]))
    return games[k], subs[k].target


def test_noalloc_solve_matches_allocating_solve_twice_in_a_row():
    big, big_target = _largest_batch_game()
    small = _std_game()
    for game, iters, target in ((big, 300, big_target), (small, 200, 0.5)):
        e0, i0, s0 = _solve(game, iters, target, noalloc=False)
        e1, i1, s1 = _solve(game, iters, target, noalloc=True)
        assert i1 == i0 and e1 == e0
        assert np.array_equal(s1, s0)
        assert i0 > 0 and s0[0] >= 0.0


def _hip_runtime():
    """The HIP runtime libpsr_dcfr.so links (already loaded: the same soname
    resolves to the same library); a stream made here is valid for it."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipStreamCreate.restype = C.c_int
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.restype = C.c_int
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return hip


def test_noalloc_solve_on_a_caller_stream():
    game = _std_game()
    e0, i0, s0 = _solve(game, 100, 0.0, noalloc=False)
    hip = _hip_runtime()
    st = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(st)) == 0
    try:
        e1, i1, s1 = _solve(game, 100, 0.0, noalloc=True, stream=st)
    finally:
        hip.hipStreamDestroy(st)
    assert (e1, i1) == (e0, i0) and np.array_equal(s1, s0)


def test_noalloc_rejects_a_game_beyond_the_array():
    game = W.turn_uniform01(1).build()   # ~17k nodes x 1128 hands: far beyond 32 MiB
    with pytest.raises(P.PsrError, match="allocation-free entry point holds at most"):
        _solve(game, 10, 0.0, noalloc=True)
