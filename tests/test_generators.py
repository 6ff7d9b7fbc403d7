"""The seeded workload generators (galax_amd/ic.py) against the definitions
of the five benchmark workloads: distributions by Kolmogorov-Smirnov
distance to their analytic CDFs, and the exact properties (ranges, masses,
truncation, the central body, per-cluster structure). The reference's own
presets (galax-init/src/lib.rs:4-122) draw other distributions, so nothing
in the reference pins these; the definitions do."""
import math

import numpy as np
import pytest

from galax_amd import ic


def _ks(samples, cdf):
    x = np.sort(samples)
    n = len(x)
    f = cdf(x)
    return float(max(np.max(np.arange(1, n + 1) / n - f), np.max(f - np.arange(n) / n)))


N = 1 << 15
KS_MAX = 1.63 / math.sqrt(N)  # 1% level


def test_uniform10k():
    b = ic.uniform()
    assert len(b["x"]) == 10_000
    for k in ("x", "y"):
        assert b[k].min() >= 0.0 and b[k].max() < 100.0
        assert _ks(b[k], lambda v: v / 100.0) < 1.63 / math.sqrt(10_000)
    assert b["m"].min() >= 0.1 and b["m"].max() < 1.0
    assert _ks(b["m"], lambda v: (v - 0.1) / 0.9) < 1.63 / math.sqrt(10_000)
    assert not b["vx"].any() and not b["vy"].any()


def test_plummer_radii_truncation_velocities():
    b = ic.plummer(N, seed=11)
    r = np.hypot(b["x"], b["y"])
    assert r.max() <= 20.0
    # r = a (u^(-2/3) - 1)^(-1/2) inverts u = F(r) = r^3 / (r^2 + a^2)^(3/2); rejection above 20a
    F = lambda v: v ** 3 / (v ** 2 + 1.0) ** 1.5
    assert _ks(r, lambda v: F(v) / F(20.0)) < KS_MAX
    th = np.arctan2(b["y"], b["x"]) % (2 * math.pi)
    assert _ks(th, lambda v: v / (2 * math.pi)) < KS_MAX
    assert np.all(b["m"] == 1.0 / N) and math.isclose(b["m"].sum(), 1.0, rel_tol=1e-12)
    # isotropic Gaussian components with sigma = sqrt(1 / (6 sqrt(1 + r^2)))
    sig = np.sqrt(1.0 / (6.0 * np.sqrt(1.0 + r ** 2)))
    for k in ("vx", "vy"):
        z = b[k] / sig
        assert abs(z.mean()) < 4.0 / math.sqrt(N) and abs(z.var() - 1.0) < 0.05
        assert _ks(z, lambda v: 0.5 * (1.0 + np.vectorize(math.erf)(v / math.sqrt(2.0)))) < KS_MAX


def test_disk_radii_and_circular_velocities():
    b = ic.disk(N, seed=12)
    R = np.hypot(b["x"], b["y"])
    assert R.max() <= 10.0
    # surface density ~ exp(-R): pdf ~ R e^(-R) on [0, 10]
    M = lambda v: (1.0 - (1.0 + v) * np.exp(-v)) / (1.0 - 11.0 * math.exp(-10.0))
    assert _ks(R, M) < KS_MAX
    assert math.isclose(b["m"].sum(), 1.0, rel_tol=1e-12) and np.all(b["m"] == b["m"][0])
    vc = np.sqrt(M(R) / R)  # enclosed mass of the truncated disk (total 1)
    tx, ty = -b["y"] / R, b["x"] / R
    vt = (b["vx"] * tx + b["vy"] * ty) / vc
    vr = (b["vx"] * b["x"] + b["vy"] * b["y"]) / (R * vc)
    assert abs(vt.mean() - 1.0) < 0.002 and abs(vt.std() - 0.05) < 0.002  # 5% dispersion
    assert abs(vr.mean()) < 0.002 and abs(vr.std() - 0.05) < 0.002


def test_galaxy_central_mass_and_test_bodies():
    n = 1 << 14
    b = ic.galaxy(n, seed=13)
    assert b["m"][0] == 1000.0 and b["x"][0] == 0.0 and b["y"][0] == 0.0
    assert b["vx"][0] == 0.0 and b["vy"][0] == 0.0
    assert np.all(b["m"][1:] == 1e-3)
    R = np.hypot(b["x"][1:], b["y"][1:])
    assert R.min() >= 1.0 and R.max() <= 50.0
    assert _ks(R, lambda v: (v - 1.0) / 49.0) < 1.63 / math.sqrt(n)  # surface density ~ 1/R
    # circular speed sqrt(M_enc(R) / R): the centre plus the test mass inside R
    order = np.argsort(R, kind="stable")
    inner = np.empty(n - 1)
    inner[order] = np.arange(n - 1)
    vc = np.sqrt((1000.0 + 1e-3 * inner) / R)
    v = np.hypot(b["vx"][1:], b["vy"][1:])
    np.testing.assert_allclose(v, vc, rtol=1e-12)
    vr = b["vx"][1:] * b["x"][1:] + b["vy"][1:] * b["y"][1:]
    assert np.abs(vr).max() <= 1e-9 * np.abs(v * R).max()  # purely tangential


def test_clustered_structure():
    n = 64 * 512
    b = ic.clustered(n, seed=14)
    assert math.isclose(b["m"].sum(), 1.0, rel_tol=1e-12)
    assert not b["vx"].any() and not b["vy"].any()
    k = n // 64
    masses, scales = [], []
    for c in range(64):
        sl = slice(c * k, (c + 1) * k)
        m = b["m"][sl]
        assert np.all(m == m[0])  # equal body masses within a cluster
        masses.append(m.sum())
        cx, cy = np.median(b["x"][sl]), np.median(b["y"][sl])
        assert -1.0 <= cx <= 101.0 and -1.0 <= cy <= 101.0  # centres ~ U[0,100]^2
        r = np.hypot(b["x"][sl] - cx, b["y"][sl] - cy)
        a_est = np.median(r) / 1.305  # median radius of the Plummer profile (u = 1/2)
        scales.append(a_est)
        assert r.max() <= 20.0 * 0.5 + 0.5  # truncated at 20 a, a <= 0.5
    lm = np.log(np.array(masses))
    assert 0.7 < lm.std() < 1.3  # log-normal(0, 1) cluster masses, normalised
    scales = np.array(scales)
    assert scales.min() > 0.04 and scales.max() < 0.6  # a ~ U[0.05, 0.5]
