"""CPU-side checks of the product library: libgalax_fmm.so loads and exports
every entry point include/galax_fmm.h declares; host-side helpers that mirror
galax-cli's parameter derivation agree with the oracle. No GPU calls."""
import os
import re

import numpy as np

from galax_amd import fmm, ic
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "galax_fmm.h")).read()
    return sorted(set(re.findall(r"GALAX_API\s+[\w\s\*]+?\b(galax_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    assert "galax_fmm_create" in names and "galax_fmm_evaluate" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    L = fmm.lib()
    for name in _declared():
        assert hasattr(L, name), name


def test_library_carries_gfx950_code_only():
    data = open(fmm.LIB_PATH, "rb").read()
    assert b"gfx950" in data
    for other in (b"gfx942", b"gfx90a", b"sm_"):
        assert other + b"\x00" not in data or other == b"sm_"


def test_version_string():
    assert "gfx950" in fmm.version()


def test_default_softening_and_dt_match_oracle():
    b = ic.plummer(3000, seed=9)
    assert fmm.default_softening(b["x"], b["y"]) == O.default_softening(b["x"], b["y"])
    ax, ay = b["vx"], b["vy"]
    assert fmm.default_dt(ax, ay) == O.default_dt(ax, ay)
    assert fmm.default_dt(np.zeros(3), np.zeros(3)) == 0.01
