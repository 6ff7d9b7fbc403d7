"""Inputs on which galax-core's build_constrained really splits leaves, and an
independent numpy re-derivation of it used to pin the oracle.

When build_constrained splits (galax-core/src/lib.rs:541-588). A leaf L of
half-width hw is split when some geometric neighbour nb (lib.rs:382-423) has
hw > 2.1 * nb.hw. For nb to contain a test point at distance 2*hw from L's
centre along an axis (within 1e-15) and still be cells_adjacent to L
(lib.rs:369-376), its centre must lie within nb.hw + 1e-15 of the point and
within hw + nb.hw + 1e-15 of L's centre, so hw <= 2*nb.hw + 2e-15. Both hold
only for nb.hw < 2e-14 (hw < ~4.2e-14): splits need leaves whose bodies
span less than about 1e-13. The constructions below therefore place bodies at
the 1e-15 scale; coincident bodies make a leaf of half-width exactly 1e-15
(lib.rs:624).

Never taken in the reference: the `split_points.contains(&mid)` and
`new_splits.contains(&mid)` tests. A leaf [s, e) contains no split point
strictly inside it (each one ends a leaf), and s < (s+e)/2 < e when e - s > 2,
so its midpoint is never already a split point; leaves of one iteration are
disjoint, so their midpoints differ.
"""
from __future__ import annotations

import numpy as np

E = 1e-15
BOUNDS = (-1.0, 1.0, -1.0, 1.0)
OFFSETS = ((2, 0), (2, 2), (0, 2), (-2, 2), (-2, 0), (-2, -2), (0, -2), (2, -2))  # lib.rs:391-398, E first, CCW


def bodies(x, y, m=None):
    n = len(x)
    z = np.zeros(n)
    return {"x": np.array(x, np.float64), "y": np.array(y, np.float64), "vx": z.copy(), "vy": z.copy(),
            "m": np.ones(n) if m is None else np.array(m, np.float64), "ax": z.copy(), "ay": z.copy()}


def hand_case():
    """Two leaves of n_max = 4 (worked by hand in test_oracle.py): a leaf of
    bodies spread over [0, 4e-15] (hw = 3e-15, centre 2e-15) and a leaf of
    coincident bodies at 6.5e-15 (hw = 1e-15). Expected leaves after
    build_constrained: [0,2), [2,6), [6,8)."""
    return bodies([0.0, 4 * E / 3, 8 * E / 3, 4 * E] + [6.5 * E] * 4, [0.0] * 8)


def random_case(seed: int, groups: int, n_max: int):
    """Groups of up to n_max bodies with random extents of 0-8e-15 along a
    random walk of 2-7e-15 steps; about one group in five is short, which
    shifts the following n_max chunks."""
    r = np.random.default_rng(seed)
    xs, ys = [], []
    cx = cy = 0.0
    for _ in range(groups):
        ext = r.choice([0.0, 0.5, 1, 2, 3, 4, 6, 8]) * E
        cx += r.uniform(2, 7) * E * r.choice([-1, 1])
        cy += r.uniform(-2, 2) * E
        k = n_max if r.uniform() < 0.8 else int(r.integers(1, n_max))
        for _ in range(k):
            xs.append(cx + r.uniform(0, ext))
            ys.append(cy + r.uniform(0, ext) * r.uniform(0, 1))
    return bodies(xs, ys, r.uniform(0.5, 1.5, len(xs)))


def mixed_case(seed: int, groups: int, n_max: int):
    """Groups of n_max/2 .. 2*n_max bodies, each coincident, alternating
    between two points 4-6e-15 apart, or uniform over 6e-15."""
    r = np.random.default_rng(seed)
    xs, ys = [], []
    cx = 0.0
    for _ in range(groups):
        kind = r.integers(0, 3)
        cx += r.uniform(3, 8) * E
        k = int(r.integers(n_max // 2, 2 * n_max))
        if kind == 0:
            xs += [cx] * k
            ys += [0.0] * k
        elif kind == 1:
            w = r.uniform(4, 6) * E
            xs += [cx + (w if i % 2 else 0.0) for i in range(k)]
            ys += [r.uniform(-1, 1) * E for _ in range(k)]
        else:
            xs += list(cx + r.uniform(0, 6 * E, k))
            ys += list(r.uniform(-1, 1, k) * E)
    return bodies(xs, ys)


def chain_case(n_max: int, seed: int = 5):
    """A leaf N of n_max coincident bodies at -2.5e-15, then a leaf L of n_max
    bodies whose first two sit at 0 and 4e-15 and the rest inside, so every
    prefix [n_max, n_max + 2^k) of L has the same box (hw = 3e-15, centre
    2e-15). N contains L's west test point (-4e-15) and is cells_adjacent to
    it, so the prefix leaf is halved on every check: with n_max = 2048 each of
    the 10 checks splits it, and the 11th build runs unchecked (lib.rs:581-587)."""
    r = np.random.default_rng(seed)
    w = 4 * E
    lx = np.concatenate([[0.0, w], r.uniform(0, w, n_max - 2)])
    ly = np.concatenate([[0.0, 0.0], r.uniform(-0.5, 0.5, n_max - 2) * E])
    return bodies([-2.5 * E] * n_max + list(lx), [0.0] * n_max + list(ly), r.uniform(0.5, 1.5, 2 * n_max))


# ------------------------------------------- independent re-derivation ----

def morton_keys(x, y, bounds):
    """t = clamp((v - lo)/(hi - lo), 0, 1 - 1e-15), u64(t * 2^21), x in the even
    bits and y in the odd bits (lib.rs:73-109), by explicit bit loops."""
    x0, x1, y0, y1 = bounds

    def q(v, lo, hi):
        t = min(max((v - lo) / (hi - lo), 0.0), 1.0 - 1e-15)
        return int(t * float(1 << 21))

    out = []
    for a, b in zip(x, y):
        xi, yi = q(float(a), x0, x1), q(float(b), y0, y1)
        k = 0
        for bit in range(21):
            k |= ((xi >> bit) & 1) << (2 * bit)
            k |= ((yi >> bit) & 1) << (2 * bit + 1)
        out.append(k)
    return out


def tree_from_ranges(x, y, m, keys, ranges):
    """Leaves from body ranges (mass, COM, bbox centre, hw = extent/2 + 1e-15)
    and parent rounds that group up to 4 consecutive nodes sharing
    key >> shift, shift = 2, 4, ... (lib.rs:602-692). Returns dict arrays."""
    cx, cy, hw, bs, be, leaf = [], [], [], [], [], []
    for s, t in ranges:
        xs, ys = x[s:t], y[s:t]
        cx.append((min(xs) + max(xs)) * 0.5)
        cy.append((min(ys) + max(ys)) * 0.5)
        hw.append(max(max(xs) - min(xs), max(ys) - min(ys)) * 0.5 + 1e-15)
        bs.append(s)
        be.append(t)
        leaf.append(1)
    lo, hi, shift = 0, len(cx), 2
    while hi - lo > 1:
        i = lo
        while i < hi:
            g = keys[bs[i]] >> shift
            j = i + 1
            while j < hi and j - i < 4 and keys[bs[j]] >> shift == g:
                j += 1
            ch = range(i, j)
            bx0 = min(cx[c] - hw[c] for c in ch)
            bx1 = max(cx[c] + hw[c] for c in ch)
            by0 = min(cy[c] - hw[c] for c in ch)
            by1 = max(cy[c] + hw[c] for c in ch)
            cx.append((bx0 + bx1) * 0.5)
            cy.append((by0 + by1) * 0.5)
            hw.append(2.0 * max(hw[c] for c in ch))
            bs.append(bs[i])
            be.append(be[j - 1])
            leaf.append(0)
            i = j
        lo, hi, shift = hi, len(cx), shift + 2
    return {"center_x": np.array(cx), "center_y": np.array(cy), "half_width": np.array(hw),
            "body_start": np.array(bs, np.uint64), "body_end": np.array(be, np.uint64),
            "is_leaf": np.array(leaf, np.uint8)}


def brute_neighbors(cx, cy, hw, i):
    """neighbors_geometric from its specification, vectorised over all nodes:
    for each test point at (+-2hw, 0, +-2hw) from node i's centre, the
    candidates are the nodes other than i whose box (half-width + 1e-15)
    contains the point and that touch node i with a size ratio <= 8; the
    nearest centre wins and equal distances go to the lowest id (a lexsort on
    (distance, id)); -1 when there is no candidate."""
    ids = np.arange(len(cx))
    ratio = np.maximum(hw, hw[i]) / np.minimum(hw, hw[i])
    adj = (ratio <= 8.0) & (np.abs(cx[i] - cx) <= hw[i] + hw + 1e-15) & (np.abs(cy[i] - cy) <= hw[i] + hw + 1e-15)
    out = []
    for a, b in OFFSETS:
        px = cx[i] + hw[i] * float(a) if a else cx[i]
        py = cy[i] + hw[i] * float(b) if b else cy[i]
        ok = (ids != i) & (np.abs(px - cx) <= hw + 1e-15) & (np.abs(py - cy) <= hw + 1e-15) & adj
        if not ok.any():
            out.append(-1)
            continue
        c = ids[ok]
        d = (px - cx[ok]) ** 2 + (py - cy[ok]) ** 2
        out.append(int(c[np.lexsort((c, d))][0]))
    return out


def rederive_constrained(b, n_max, bounds=BOUNDS):
    """build_constrained (lib.rs:541-588) re-derived: leaves of the current
    order cut at n_max and at the split points; up to 10 checks, each
    splitting every leaf of more than 2 bodies that has a neighbour of less
    than 1/2.1 its half-width at (start+end)//2; then one more build.
    Returns (tree dict, stats) with stats = iterations (checks run), splits,
    and `small` (violating leaves of <= 2 bodies, left unsplit)."""
    x, y, m = list(b["x"]), list(b["y"]), list(b["m"])
    n = len(x)
    keys = morton_keys(x, y, bounds)
    splits: list[int] = []
    stats = {"iterations": 0, "splits": 0, "small": 0}
    for it in range(11):
        ranges, s = [], 0
        while s < n:
            bound = min(s + n_max, n)
            inside = [p for p in splits if s < p < bound]
            t = min(inside) if inside else bound
            ranges.append((s, t))
            s = t
        tree = tree_from_ranges(x, y, m, keys, ranges)
        if it == 10:
            break
        stats["iterations"] = it + 1
        cx, cy, hw = tree["center_x"], tree["center_y"], tree["half_width"]
        new = []
        for lid in np.nonzero(tree["is_leaf"])[0]:
            for nb in brute_neighbors(cx, cy, hw, int(lid)):
                if nb >= 0 and hw[lid] > hw[nb] * 2.1:
                    s0, e0 = int(tree["body_start"][lid]), int(tree["body_end"][lid])
                    if e0 - s0 > 2:
                        new.append((s0 + e0) // 2)
                    else:
                        stats["small"] += 1
                    break
        stats["splits"] += len(new)
        if not new:
            break
        splits = sorted(set(splits) | set(new))
    return tree, stats


# (name, generator, n_max): every case splits at least once in the oracle.
SPLIT_CASES = [
    ("hand", hand_case, 4),
    ("random35", lambda: random_case(35, 16, 8), 8),
    ("random37", lambda: random_case(37, 16, 8), 8),
    ("random6", lambda: random_case(6, 16, 4), 4),
    ("mixed306", lambda: mixed_case(306, 24, 8), 8),
    ("mixed71", lambda: mixed_case(71, 24, 32), 32),
    ("chain256", lambda: chain_case(256), 256),
    ("chain2048_cap", lambda: chain_case(2048), 2048),
]
