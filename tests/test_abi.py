"""The C-ABI library loads and exports every entry point include/g5hip.h
declares (no device work: runs without a GPU)."""
import os
import re

import pytest

import gpugraph500_amd as g5

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "g5hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(g5_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "g5_bfs_expand" in syms and "g5_s4_encode" in syms and len(syms) >= 25


def test_library_exports_every_declared_symbol():
    if not os.path.exists(g5.LIB_PATH):
        pytest.fail(f"{g5.LIB_PATH} not built (run python -m gpugraph500_amd.build)")
    L = g5.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_s4_max_words_bound_matches_oracle():
    import oracle_lib as O
    L = g5.lib()
    for n in (0, 1, 127, 128, 129, 2047, 2048, 2049, 10 ** 6 + 3):
        assert L.g5_s4_max_words(n) == O.lib().orc_s4bp128d4_max_words(O.c_sz(n))


def test_last_error_is_a_string():
    assert isinstance(g5.lib().g5_last_error(), bytes)
