"""galax_direct_sum_noalloc (no counterpart in the reference; for callers that
cannot allocate device memory) against galax_direct_sum, through the C ABI:
bit-identical accelerations and potentials on two calls in a row, the second
with a smaller input, and sizes above the scratch maximum rejected on the
host."""
import ctypes as C

import numpy as np
import pytest

from galax_amd import fmm, ic

pytestmark = pytest.mark.gpu


def _lib():
    L = fmm.lib()
    vp, sz, d = C.c_void_p, C.c_size_t, C.c_double
    for name in ("galax_direct_sum", "galax_direct_sum_noalloc"):
        f = getattr(L, name)
        f.argtypes = [vp, vp, vp, sz, d, vp, sz, vp, vp, vp]
        f.restype = C.c_int
    L.galax_direct_sum_noalloc_bytes.argtypes = [sz, sz]
    L.galax_direct_sum_noalloc_bytes.restype = sz
    return L


def _call(L, name, b, eps, targets):
    n = len(b["x"])
    nt = len(targets) if targets is not None else n
    ax, ay, phi = np.zeros(nt), np.zeros(nt), np.zeros(nt)
    t = None if targets is None else np.ascontiguousarray(targets, np.uint64)
    rc = getattr(L, name)(b["x"].ctypes.data, b["y"].ctypes.data, b["m"].ctypes.data, n, eps,
                          None if t is None else t.ctypes.data, nt, ax.ctypes.data, ay.ctypes.data, phi.ctypes.data)
    assert rc == 0, (name, rc, L.galax_fmm_last_error())
    return ax, ay, phi


def test_direct_sum_noalloc_matches_allocating_entry_point():
    L = _lib()
    cases = [
        ("plummer 2^20, 10,000 sampled targets", ic.plummer(1 << 20, seed=501), 0.05,
         (ic.SplitMix64(7).next_u64(10_000) % np.uint64(1 << 20))),
        ("plummer 60,000, every body a target (smaller)", ic.plummer(60_000, seed=502), 0.02, None),
    ]
    for label, b, eps, targets in cases:
        print(f"no-alloc direct sum: {label}", flush=True)
        n = len(b["x"])
        nt = n if targets is None else len(targets)
        assert L.galax_direct_sum_noalloc_bytes(n, nt) <= 64 << 20
        want = _call(L, "galax_direct_sum", b, eps, targets)
        got = _call(L, "galax_direct_sum_noalloc", b, eps, targets)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
        assert np.isfinite(want[0]).all() and np.abs(want[0]).max() > 0


def test_direct_sum_noalloc_rejects_sizes_above_its_scratch():
    L = _lib()
    n = 3_000_000  # 3 x 24 MB of staging alone exceeds 64 MiB
    assert L.galax_direct_sum_noalloc_bytes(n, 10) > 64 << 20
    x = np.zeros(n)
    ax, ay = np.zeros(10), np.zeros(10)
    t = np.arange(10, dtype=np.uint64)
    rc = L.galax_direct_sum_noalloc(x.ctypes.data, x.ctypes.data, x.ctypes.data, n, 0.1, t.ctypes.data, 10,
                                    ax.ctypes.data, ay.ctypes.data, None)
    assert rc == -4
    assert b"maximum" in L.galax_fmm_last_error()
