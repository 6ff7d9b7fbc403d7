"""Seeded synthetic initial conditions for the five benchmark workloads.

The reference's own presets (galax-init/src/lib.rs:4-122) use unseeded
`fastrand` and different distributions, so these generators are new: they draw
from SplitMix64 (Steele, Lea & Flood 2014) turned into 53-bit uniform doubles,
which makes every workload reproducible bit for bit on any host.

Workloads (N, p, n_max, bounds):
  uniform10k   N=10,000   x,y ~ U[0,100], m ~ U[0.1,1.0], v = 0; bounds [0,100]^2
  plummer1m    N=2^20     2D Plummer, a=1, truncated at 20a, m=1/N, Gaussian v
  disk4m       N=2^22     exponential disk, scale 1, truncated at 10, m=1/N
  galaxy262k   N=2^18     central mass 1000 + test bodies of mass 1e-3, R~U[1,50]
  clustered16m N=2^24     64 Plummer clusters, log-normal cluster masses
"""
from __future__ import annotations

import math

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """SplitMix64 stream; uniform() returns doubles in [0, 1) with 53 random bits."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * _GAMMA
            self.state = self.state + np.uint64(n) * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int) -> np.ndarray:
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller on pairs of uniforms (1-u keeps the log finite)."""
        u1 = 1.0 - self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def _bodies(n):
    return {k: np.zeros(n) for k in ("x", "y", "vx", "vy", "m", "ax", "ay")}


def uniform(n=10_000, seed=20260001, lo=0.0, hi=100.0):
    """x, y ~ U[lo,hi]^2, m ~ U[0.1, 1.0], zero velocities."""
    r = SplitMix64(seed)
    b = _bodies(n)
    b["x"][:] = lo + (hi - lo) * r.uniform(n)
    b["y"][:] = lo + (hi - lo) * r.uniform(n)
    b["m"][:] = 0.1 + 0.9 * r.uniform(n)
    return b


def _plummer_radii(r: SplitMix64, n: int, a: float, rmax: float) -> np.ndarray:
    out = np.empty(0)
    while out.size < n:
        u = r.uniform(int((n - out.size) * 1.05) + 64)
        u = u[u > 0.0]
        rad = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        out = np.concatenate([out, rad[rad <= rmax]])
    return out[:n]


def plummer(n=1 << 20, seed=20260002, a=1.0, rmax_over_a=20.0):
    """2D Plummer: r = a (u^(-2/3) - 1)^(-1/2), rejected above 20a, angle
    uniform, isotropic Gaussian velocities with sigma = sqrt(1/(6 sqrt(1+r^2/a^2)))."""
    r = SplitMix64(seed)
    b = _bodies(n)
    rad = _plummer_radii(r, n, a, rmax_over_a * a)
    th = 2.0 * math.pi * r.uniform(n)
    b["x"][:] = rad * np.cos(th)
    b["y"][:] = rad * np.sin(th)
    sig = np.sqrt(1.0 / (6.0 * np.sqrt(1.0 + (rad / a) ** 2)))
    b["vx"][:] = sig * r.normal(n)
    b["vy"][:] = sig * r.normal(n)
    b["m"][:] = 1.0 / n
    return b


def disk(n=1 << 22, seed=20260003, scale=1.0, rmax=10.0, disp=0.05):
    """Exponential disk: surface density ~ exp(-R/scale) (pdf ~ R e^(-R)),
    truncated at rmax, equal masses summing to 1, circular velocities from the
    enclosed mass of the truncated disk with a 5% Gaussian dispersion."""
    r = SplitMix64(seed)
    b = _bodies(n)
    out = np.empty(0)
    while out.size < n:
        k = int((n - out.size) * 1.05) + 64
        u1 = 1.0 - r.uniform(k)
        u2 = 1.0 - r.uniform(k)
        rad = -scale * np.log(u1 * u2)  # Gamma(2, scale)
        out = np.concatenate([out, rad[rad <= rmax]])
    rad = out[:n]
    th = 2.0 * math.pi * r.uniform(n)
    b["x"][:] = rad * np.cos(th)
    b["y"][:] = rad * np.sin(th)
    cdf = lambda R: 1.0 - (1.0 + R / scale) * np.exp(-R / scale)
    menc = cdf(rad) / cdf(rmax)
    vc = np.sqrt(menc / np.maximum(rad, 1e-12))
    vt = vc * (1.0 + disp * r.normal(n))
    vr = vc * disp * r.normal(n)
    b["vx"][:] = -vt * np.sin(th) + vr * np.cos(th)
    b["vy"][:] = vt * np.cos(th) + vr * np.sin(th)
    b["m"][:] = 1.0 / n
    return b


def galaxy(n=1 << 18, seed=20260004, m_center=1000.0, m_test=1e-3, r_in=1.0, r_out=50.0):
    """Body 0: mass 1000 at rest at the origin. Bodies 1..n-1: mass 1e-3,
    R ~ U[r_in, r_out] (surface density ~ 1/R), uniform angle, circular speed
    sqrt(M_enc(R)/R) where M_enc counts the centre AND the test mass inside R."""
    r = SplitMix64(seed)
    b = _bodies(n)
    k = n - 1
    rad = r_in + (r_out - r_in) * r.uniform(k)
    th = 2.0 * math.pi * r.uniform(k)
    order = np.argsort(rad, kind="stable")
    inner = np.empty(k)
    inner[order] = np.arange(k)
    menc = m_center + m_test * inner
    vc = np.sqrt(menc / rad)
    b["x"][1:] = rad * np.cos(th)
    b["y"][1:] = rad * np.sin(th)
    b["vx"][1:] = -vc * np.sin(th)
    b["vy"][1:] = vc * np.cos(th)
    b["m"][0] = m_center
    b["m"][1:] = m_test
    return b


def clustered(n=1 << 24, seed=20260005, n_clusters=64):
    """64 Plummer clusters, equal body counts (n/64 each), cluster masses
    log-normal(0,1) normalised to 1 (equal body masses within a cluster),
    scale radii ~ U[0.05,0.5], centres ~ U[0,100]^2, truncated at 20a,
    zero velocities."""
    r = SplitMix64(seed)
    b = _bodies(n)
    cm = np.exp(r.normal(n_clusters))
    cm /= cm.sum()
    ca = 0.05 + 0.45 * r.uniform(n_clusters)
    cx = 100.0 * r.uniform(n_clusters)
    cy = 100.0 * r.uniform(n_clusters)
    counts = np.full(n_clusters, n // n_clusters)
    counts[: n - counts.sum()] += 1
    s = 0
    for c in range(n_clusters):
        k = int(counts[c])
        rad = _plummer_radii(r, k, ca[c], 20.0 * ca[c])
        th = 2.0 * math.pi * r.uniform(k)
        b["x"][s:s + k] = cx[c] + rad * np.cos(th)
        b["y"][s:s + k] = cy[c] + rad * np.sin(th)
        b["m"][s:s + k] = cm[c] / k
        s += k
    return b


# Workload table: generator, p values, n_max, bounds ("data" = data bbox).
WORKLOADS = {
    "uniform10k": dict(gen=uniform, n=10_000, p=(4,), n_max=32, bounds=(0.0, 100.0, 0.0, 100.0)),
    "plummer1m": dict(gen=plummer, n=1 << 20, p=(4, 6), n_max=32, bounds=(-50.0, 50.0, -50.0, 50.0)),
    "disk4m": dict(gen=disk, n=1 << 22, p=(2, 4, 6, 8), n_max=32, bounds=(-50.0, 50.0, -50.0, 50.0)),
    "galaxy262k": dict(gen=galaxy, n=1 << 18, p=(6,), n_max=32, bounds=(-50.0, 50.0, -50.0, 50.0)),
    "clustered16m": dict(gen=clustered, n=1 << 24, p=(8,), n_max=64, bounds="data"),
}


def make(name: str, n: int | None = None, seed: int | None = None):
    """Returns (bodies, spec) for a named workload, optionally at a smaller n."""
    spec = dict(WORKLOADS[name])
    kw = {}
    if n is not None:
        kw["n"] = n
    if seed is not None:
        kw["seed"] = seed
    b = spec["gen"](**kw)
    if spec["bounds"] == "data":
        spec["bounds"] = (float(b["x"].min()), float(b["x"].max()), float(b["y"].min()), float(b["y"].max()))
    return b, spec


def reference_test_bodies(n: int, x0=5.0, span=90.0):
    """The reference's RNG-free test distribution (galax-core/src/lib.rs:2320-2329):
    x = x0 + span*frac(3.7 t), y = x0 + span*frac(7.3 t), m = 0.1 + 10*frac(11 t), t = i/n."""
    b = _bodies(n)
    for i in range(n):
        t = i / n
        b["x"][i] = x0 + span * math.modf(t * 3.7)[0]
        b["y"][i] = x0 + span * math.modf(t * 7.3)[0]
        b["m"][i] = 0.1 + 10.0 * math.modf(t * 11.0)[0]
    return b
