"""MI355X range-solver DCFR path: Python mirror of the reference's
range-solver / gpu-range-solver API over the libpsr_dcfr.so C ABI.

Names follow crates/range-solver and crates/gpu-range-solver:
``card_pair_to_index``, ``Range``, ``TreeConfig``, ``ActionTree``,
``CardConfig``, ``PostFlopGame.with_config``, ``GpuSolverConfig``,
``gpu_solve_hand_parallel``, ``GpuBatchSolver``.  Every solve runs on the GPU
through the HIP kernels; there is no CPU fallback (the library raises when
the device path is unavailable).
"""
from .range_solver import (  # noqa: F401
    NOT_DEALT,
    PLAYER_CHANCE,
    PLAYER_CHANCE_FLAG,
    PLAYER_DEPTH_BOUNDARY_FLAG,
    PLAYER_FOLD_FLAG,
    PLAYER_IP,
    PLAYER_MASK,
    PLAYER_OOP,
    PLAYER_TERMINAL_FLAG,
    Action,
    ActionTree,
    BoardState,
    CardConfig,
    GpuBatchSolver,
    GpuSolveResult,
    GpuSolverConfig,
    PostFlopGame,
    PsrError,
    Range,
    TreeConfig,
    action_to_string,
    card_from_str,
    card_to_string,
    card_pair_to_index,
    flop_from_str,
    gpu_solve_hand_parallel,
    hand_evaluate,
    index_to_card_pair,
    lib_path,
    load_library,
)
