"""The C ABI library loads and exports every entry point include/psr_dcfr.h
declares (no device calls)."""
import ctypes
import os
import re

import poker_solver_rust_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "psr_dcfr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "psr_batch_solve" in names and "psr_game_build" in names and len(names) > 30


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(P.lib_path())
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_host_entry_points_work_without_device():
    assert P.card_pair_to_index(0, 1) == 0
    assert P.load_library().psr_last_error() is not None
