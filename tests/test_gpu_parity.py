"""Parity of the HIP path (libgalax_fmm.so through its C ABI) with the CPU
oracle, on the GPU: bit-exact tree, permutation and lists; forces within the
north-star tolerance (median relative difference <= 1e-6 in FP64); the
reference's known-answer tests; edge cases."""
import os

import numpy as np
import pytest

from galax_amd import fmm, ic
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

TREE_FIELDS = ("center_x", "center_y", "half_width", "mass", "com_x", "com_y", "body_start", "body_end",
               "children", "parent", "is_leaf")


def _bodies(d):
    return fmm.BodiesSoA.from_dict(d)


def _ctx(n, p=4, n_max=32, eps=0.1, bounds=(0.0, 100.0, 0.0, 100.0), **kw):
    return fmm.GpuContext(fmm.GpuConfig(max_n=max(n, 1), p=p, eps=eps, n_max=n_max, bounds=bounds, **kw))


def _gpu_build(d, **kw):
    n = len(d["x"])
    ctx = _ctx(n, **kw)
    b = _bodies(d)
    perm = ctx.build_tree(b)
    return ctx, b, perm


def _oracle_build(d, n_max, bounds, constrained=False):
    od = {k: np.array(d[k], np.float64) for k in ("x", "y", "vx", "vy", "m", "ax", "ay")}
    if constrained:
        t = O.build_constrained(od, n_max, bounds)
        perm = np.arange(len(od["x"]), dtype=np.uint64)
    else:
        t, perm = O.tree_build(od, n_max, bounds, return_perm=True)
    return t, od, perm


def _assert_tree_equal(gt, ot):
    for k in TREE_FIELDS:
        a, b = np.asarray(gt[k]), np.asarray(ot[k])
        assert a.shape == b.shape, k
        if a.dtype == np.float64:
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), f"{k} differs at {np.nonzero(a != b)[0][:5]}"
        else:
            assert np.array_equal(a.astype(np.int64), b.astype(np.int64)), f"{k} differs"


def _assert_lists_equal(g_csr, o_lists):
    go, gd = g_csr
    oo, od = o_lists.csr()
    go, oo = go.astype(np.int64), oo.astype(np.int64)
    if not np.array_equal(go, oo):
        bad = np.nonzero(go != oo)[0]
        raise AssertionError(f"list offsets differ first at node {bad[0]}: {go[bad[0]-1:bad[0]+2]} vs "
                             f"{oo[bad[0]-1:bad[0]+2]} (totals {go[-1]} vs {oo[-1]})")
    gd, od = gd.astype(np.int64), od.astype(np.int64)
    if not np.array_equal(gd, od):
        bad = np.nonzero(gd != od)[0]
        raise AssertionError(f"list entries differ at {len(bad)} places, first {bad[0]}: {gd[bad[0]]} vs {od[bad[0]]}")


def _snap(name):
    """Bodies of one of the reference's snapshots (post-step states of its own
    runs, runs/*/snap_*.bin, copied under tests/golden/)."""
    b = fmm.read_snapshot(os.path.join(GOLDEN, name))
    return {"x": b.x, "y": b.y, "vx": b.vx, "vy": b.vy, "m": b.mass, "ax": b.ax, "ay": b.ay}


CASES = [
    ("uniform", lambda: ic.uniform(6000, seed=101), 32, (0.0, 100.0, 0.0, 100.0)),
    ("snap_demo", lambda: _snap("galx_demo_0010.bin"), 32, (-50.0, 50.0, -50.0, 50.0)),
    ("snap_test", lambda: _snap("galx_test_0005.bin"), 8, (-50.0, 50.0, -50.0, 50.0)),
    ("snap_viz_demo", lambda: _snap("galx_viz_demo_0010.bin"), 16, (-50.0, 50.0, -50.0, 50.0)),
    ("plummer", lambda: ic.plummer(8000, seed=102), 32, (-50.0, 50.0, -50.0, 50.0)),
    ("clamped_dupes", lambda: ic.uniform(5000, seed=103), 32, (-50.0, 50.0, -50.0, 50.0)),  # 75% clamp (CLI bounds)
    ("galaxy", lambda: ic.galaxy(4096, seed=104), 32, (-50.0, 50.0, -50.0, 50.0)),
    ("tiny_leaves", lambda: ic.uniform(700, seed=105), 1, (0.0, 100.0, 0.0, 100.0)),
    ("n_max64", lambda: ic.clustered(8192, seed=106, n_clusters=8), 64, "data"),
]


def _bounds(d, bounds):
    if bounds == "data":
        return (float(d["x"].min()), float(d["x"].max()), float(d["y"].min()), float(d["y"].max()))
    return bounds


@pytest.mark.parametrize("name,gen,n_max,bounds", CASES, ids=[c[0] for c in CASES])
def test_tree_and_lists_bit_exact(name, gen, n_max, bounds):
    d = gen()
    bounds = _bounds(d, bounds)
    ctx, b, perm = _gpu_build(d, n_max=n_max, bounds=bounds)
    ot, od, operm = _oracle_build(d, n_max, bounds)
    assert np.array_equal(perm, operm)
    assert np.array_equal(b.x.view(np.uint64), od["x"].view(np.uint64))
    assert np.array_equal(b.mass, od["m"])
    gt, gm2l, gp2p = ctx.export_tree()
    _assert_tree_equal(gt, ot.export())
    _assert_lists_equal(gm2l, O.m2l_lists(ot))
    _assert_lists_equal(gp2p, O.p2p_lists(ot))


@pytest.mark.parametrize("prec", [fmm.PREC_FP64, fmm.PREC_FP64_EXACT, fmm.PREC_MIXED_FP32])
def test_upper_level_m2l_rows_built_on_demand(prec):
    """build_tree walks the M2L rows of levels 0 and 1 only, the rows an
    evaluation reads; the others are built when the lists or their total are
    asked for. The forces before and after are equal bit for bit, the
    reported reference pair count is 0 until then, and the completed lists
    equal the oracle's."""
    d = ic.plummer(12_000, seed=91)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    ctx, b, _ = _gpu_build(d, p=4, eps=0.05, bounds=bounds, precision=prec)
    ctx.evaluate(b)
    first = (b.ax.copy(), b.ay.copy())
    assert ctx.timings().m2l_pairs_reference == 0
    info = ctx.tree_info()
    assert info["n_levels"] > 2
    gt, gm2l, gp2p = ctx.export_tree()
    assert gm2l[1].size == info["m2l_total"]
    ctx.evaluate(b)
    assert np.array_equal(b.ax, first[0]) and np.array_equal(b.ay, first[1])
    assert ctx.timings().m2l_pairs_reference == info["m2l_total"]
    ot, _, _ = _oracle_build(d, 32, bounds)
    _assert_lists_equal(gm2l, O.m2l_lists(ot))
    _assert_lists_equal(gp2p, O.p2p_lists(ot))


def test_tree_bit_exact_full_size_plummer1m():
    """Tree and permutation at the 2^20-body Plummer workload (lists are
    checked at the smaller sizes above: the oracle's list builder is O(n^2))."""
    d = ic.plummer(1 << 20)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    ctx, b, perm = _gpu_build(d, n_max=32, bounds=bounds)
    ot, od, operm = _oracle_build(d, 32, bounds)
    assert np.array_equal(perm, operm)
    gt, _, _ = ctx.export_tree()
    _assert_tree_equal(gt, ot.export())


def test_constrained_tree_bit_exact():
    d = ic.reference_test_bodies(500)
    bounds = (0.0, 100.0, 0.0, 100.0)
    ctx, b, perm = _gpu_build(d, n_max=16, bounds=bounds, tree_mode=fmm.TREE_CONSTRAINED)
    ot, od, _ = _oracle_build(d, 16, bounds, constrained=True)
    gt, gm2l, gp2p = ctx.export_tree()
    _assert_tree_equal(gt, ot.export())
    _assert_lists_equal(gm2l, O.m2l_lists(ot))
    _assert_lists_equal(gp2p, O.p2p_lists(ot))


def _split_cases():
    from constrained_cases import SPLIT_CASES
    return SPLIT_CASES


@pytest.mark.parametrize("name,gen,n_max", _split_cases(), ids=[c[0] for c in _split_cases()])
def test_constrained_tree_splits_bit_exact(name, gen, n_max):
    """build_constrained on inputs where it really splits leaves (bodies at the
    1e-15 scale; chain2048_cap splits on all 10 checks): the device's
    neighbour search and split loop give the oracle's tree and lists bit for
    bit, with more leaves than plain n_max chunks (galax-core/src/lib.rs:382-423,
    541-588)."""
    from constrained_cases import BOUNDS

    d = gen()
    n = len(d["x"])
    print(f"constrained split case {name}: n={n} n_max={n_max}", flush=True)
    ctx, b, perm = _gpu_build(d, n_max=n_max, bounds=BOUNDS, tree_mode=fmm.TREE_CONSTRAINED)
    assert np.array_equal(perm, np.arange(n))  # build_constrained does not sort
    ot, od, _ = _oracle_build(d, n_max, BOUNDS, constrained=True)
    gt, gm2l, gp2p = ctx.export_tree()
    _assert_tree_equal(gt, ot.export())
    assert int(np.asarray(gt["is_leaf"]).sum()) > -(-n // n_max)
    _assert_lists_equal(gm2l, O.m2l_lists(ot))
    _assert_lists_equal(gp2p, O.p2p_lists(ot))
    if n_max <= 32:
        # forces on the split tree (FP64, p = 4, eps = 0.1): near the oracle's
        ctx.evaluate(b)
        ax, ay, _ = O.compute_fmm_force(od["x"], od["y"], od["m"], ot, O.m2l_lists(ot), O.p2p_lists(ot), 4, 0.1)
        rd = np.hypot(b.ax - ax, b.ay - ay) / np.maximum(np.hypot(ax, ay), 1e-300)
        assert float(np.median(rd)) <= 1e-9 and float(rd.max()) <= 1e-6


def _rel_diff(a, b):
    den = np.maximum(np.hypot(b[0], b[1]), 1e-300)
    return np.hypot(a[0] - b[0], a[1] - b[1]) / den


@pytest.mark.parametrize("prec", [fmm.PREC_FP64, fmm.PREC_FP64_EXACT])
@pytest.mark.parametrize("p", [0, 2, 4, 6, 8])
def test_forces_match_oracle_reference(p, prec):
    d = ic.plummer(12000, seed=201)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    eps = 0.02
    ctx, b, _ = _gpu_build(d, p=p, eps=eps, bounds=bounds, precision=prec)
    ctx.evaluate(b)
    ot, od, _ = _oracle_build(d, 32, bounds)
    ax, ay, phi = O.compute_fmm_force(od["x"], od["y"], od["m"], ot, O.m2l_lists(ot), O.p2p_lists(ot), p, eps,
                                      want_phi=True)
    rd = _rel_diff((b.ax, b.ay), (ax, ay))
    assert np.median(rd) <= 1e-6, np.median(rd)
    assert np.max(rd) <= 1e-9, np.max(rd)
    np.testing.assert_allclose(b.phi, phi, rtol=1e-9, atol=1e-12 * np.abs(phi).max())


def _levels(ex):
    """Node levels of an exported tree (ids are level-ordered, children first)."""
    lv = np.zeros(len(ex["is_leaf"]), np.int64)
    for i in range(len(lv)):
        if not ex["is_leaf"][i]:
            lv[i] = 1 + lv[ex["children"][i][0]]
    return lv


@pytest.mark.parametrize("p", [2, 4, 6, 8])
def test_expansions_bit_exact(p):
    """P2M + M2M (levels 0-1) bit for bit in both FP64 forms; the M2L locals of
    levels 0-1 bit for bit in FP64_EXACT (galax-core's per-pair order) and to
    rounding in FP64 (matrix cores, grouped by key)."""
    d = ic.plummer(12000, seed=206)
    bounds, eps = (-50.0, 50.0, -50.0, 50.0), 0.02
    ot, od, _ = _oracle_build(d, 32, bounds)
    ex = ot.export()
    live = _levels(ex) <= 1
    S = (p + 1) * (p + 2) // 2
    mp_o = O.p2m(od["x"], od["y"], od["m"], ot, p)
    O.m2m(mp_o, ot, p)
    loc_o = np.zeros_like(mp_o)
    O.m2l(loc_o, mp_o, O.m2l_lists(ot), ot, p, eps)
    mp_o, loc_o = mp_o.reshape(-1, S)[live], loc_o.reshape(-1, S)[live]
    for prec in (fmm.PREC_FP64_EXACT, fmm.PREC_FP64):
        ctx, b, _ = _gpu_build(d, p=p, eps=eps, bounds=bounds, precision=prec)
        ctx.evaluate(b)
        mp_g, loc_g = ctx.export_expansions()
        mp_g, loc_g = mp_g[live], loc_g[live]
        assert np.array_equal(mp_g.view(np.uint64), mp_o.view(np.uint64)), (prec, np.argwhere(mp_g != mp_o)[:5])
        if prec == fmm.PREC_FP64_EXACT:
            assert np.array_equal(loc_g.view(np.uint64), loc_o.view(np.uint64)), np.argwhere(loc_g != loc_o)[:5]
        else:
            scale = np.abs(loc_o).max(axis=1, keepdims=True)
            assert np.max(np.abs(loc_g - loc_o) / np.maximum(scale, 1e-300)) < 1e-12


@pytest.mark.parametrize("name", ["galx_demo_0010.bin", "galx_viz_demo_0010.bin"])
def test_forces_on_reference_snapshots(name):
    """Forces on post-step states of the reference's own runs (GALX fixtures)."""
    d = _snap(name)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    eps, p = 0.05, 4
    ctx, b, _ = _gpu_build(d, p=p, eps=eps, bounds=bounds)
    ctx.evaluate(b)
    ot, od, _ = _oracle_build(d, 32, bounds)
    ax, ay, _ = O.compute_fmm_force(od["x"], od["y"], od["m"], ot, O.m2l_lists(ot), O.p2p_lists(ot), p, eps,
                                    want_phi=True)
    rd = _rel_diff((b.ax, b.ay), (ax, ay))
    assert np.median(rd) <= 1e-6, np.median(rd)
    assert np.max(rd) <= 1e-9, np.max(rd)


def test_uniform10k_forces_and_validate():
    """The 10,000-body uniform workload at p=4: forces vs the oracle, and the
    --validate metric vs the oracle's on the same samples within 5%."""
    d = ic.uniform(10_000)
    bounds = (0.0, 100.0, 0.0, 100.0)
    ot, od, _ = _oracle_build(d, 32, bounds)
    eps = O.default_softening(od["x"], od["y"])
    ctx, b, _ = _gpu_build(d, p=4, eps=eps, bounds=bounds)
    ctx.evaluate(b)
    m2l, p2p = O.m2l_lists(ot), O.p2p_lists(ot)
    ax, ay, _ = O.compute_fmm_force(od["x"], od["y"], od["m"], ot, m2l, p2p, 4, eps)
    rd = _rel_diff((b.ax, b.ay), (ax, ay))
    assert np.median(rd) <= 1e-6
    samples = np.arange(0, 10_000, 10, dtype=np.uint64)
    mx, mean, _ = ctx.validate(b, samples)
    ex, ey, _ = O.n2_force(od["x"], od["y"], od["m"], eps, targets=samples)
    fa = np.hypot(ax[samples.astype(int)], ay[samples.astype(int)])
    ea = np.hypot(ex, ey)
    rel = np.abs(fa - ea) / np.maximum(ea, 1e-20)
    assert abs(mx - rel.max()) <= 0.05 * rel.max()
    assert abs(mean - rel.mean()) <= 0.05 * rel.mean()


def test_p0_equals_p2p_only():  # galax-core/src/lib.rs:2320-2357
    d = ic.reference_test_bodies(200)
    bounds = (0.0, 100.0, 0.0, 100.0)
    ctx, b, _ = _gpu_build(d, p=0, n_max=16, eps=0.1, bounds=bounds)
    ctx.evaluate(b)
    ot, od, _ = _oracle_build(d, 16, bounds)
    px, py = O.p2p(od["x"], od["y"], od["m"], ot, O.p2p_lists(ot), 0.1)
    assert np.max(np.abs(b.ax - px)) < 1e-14 and np.max(np.abs(b.ay - py)) < 1e-14


def test_upload_tree_matches_device_build():
    d = ic.plummer(5000, seed=202)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    ot, od, _ = _oracle_build(d, 32, bounds)
    ctx = _ctx(5000, p=6, eps=0.05, bounds=bounds)
    ctx.upload_tree_data(ot.export(), O.m2l_lists(ot).csr(), O.p2p_lists(ot).csr())
    b = _bodies(od)
    ctx.evaluate(b)
    ax, ay, _ = O.compute_fmm_force(od["x"], od["y"], od["m"], ot, O.m2l_lists(ot), O.p2p_lists(ot), 6, 0.05)
    assert np.median(_rel_diff((b.ax, b.ay), (ax, ay))) <= 1e-6


def _swap_nodes(tree, lists, p, q):
    """The same tree with nodes p and q exchanged: every per-node array
    reordered, every stored node id (children, parent, list entries) mapped."""
    nn = len(tree["center_x"])
    perm = np.arange(nn)
    perm[p], perm[q] = q, p  # an involution: new node j is old node perm[j], old id i is new id perm[i]
    new = {k: np.asarray(v)[perm].copy() for k, v in tree.items()}
    for k in ("children", "parent"):
        a = new[k]
        a[a >= 0] = perm[a[a >= 0]]
    out = []
    for off, dat in lists:
        off = np.asarray(off, np.int64)
        rows = [np.asarray(dat[off[perm[j]]:off[perm[j] + 1]], np.int64) for j in range(nn)]
        noff = np.zeros(nn + 1, np.uint64)
        noff[1:] = np.cumsum([len(r) for r in rows])
        ndat = perm[np.concatenate(rows)].astype(np.uint32) if noff[-1] else np.zeros(0, np.uint32)
        out.append((noff, ndat))
    return new, out


def test_corrected_one_child_levels_not_aligned():
    """An uploaded tree whose one-child levels are not aligned with the level
    below (two parents of such a level exchanged): those levels take the
    per-level shifts instead of the fused chains (k_chain_check), and the
    corrected variant's forces equal those on the same tree in its built
    order to rounding."""
    d = ic.plummer(20_000, seed=203)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    ot, od, _ = _oracle_build(d, 32, bounds)
    tree, lv = ot.export(), ot.levels()
    off = np.searchsorted(lv, np.arange(lv.max() + 2))
    sizes = np.diff(off)
    assert sizes[3] == sizes[2] == sizes[4] and sizes[3] > 1  # levels 1..: one parent per leaf here
    lists = [O.m2l_lists(ot).csr(), O.p2p_lists(ot).csr()]
    res = []
    for t, (m2l, p2p) in ((tree, lists), _swap_nodes(tree, lists, int(off[3]), int(off[3]) + 1)):
        ctx = _ctx(len(od["x"]), p=6, eps=0.05, bounds=bounds, algo=fmm.ALGO_CORRECTED, theta=0.4)
        ctx.upload_tree_data(t, m2l, p2p)
        b = _bodies(od)
        ctx.evaluate(b)
        res.append((b.ax.copy(), b.ay.copy()))
    rd = _rel_diff(res[1], res[0])
    assert np.isfinite(rd).all() and rd.max() <= 1e-10, rd.max()


CORRECTED_LIST_CASES = [
    ("plummer_t03", lambda: ic.plummer(20000, seed=203), 32, (-50.0, 50.0, -50.0, 50.0), 0.3),
    ("uniform_t05", lambda: ic.uniform(6000, seed=101), 32, (0.0, 100.0, 0.0, 100.0), 0.5),
    ("snap_demo_t04", lambda: _snap("galx_demo_0010.bin"), 32, (-50.0, 50.0, -50.0, 50.0), 0.4),
    ("tiny_leaves_t05", lambda: ic.uniform(700, seed=105), 1, (0.0, 100.0, 0.0, 100.0), 0.5),
    ("galaxy_t03", lambda: ic.galaxy(4096, seed=104), 16, (-50.0, 50.0, -50.0, 50.0), 0.3),
]


@pytest.mark.parametrize("name,gen,n_max,bounds,theta", CORRECTED_LIST_CASES, ids=[c[0] for c in CORRECTED_LIST_CASES])
def test_corrected_lists_bit_exact(name, gen, n_max, bounds, theta):
    """The device dual tree traversal gives the oracle's corrected-variant
    lists entry for entry (M2L ascending; P2P the leaf first, then ascending)."""
    d = gen()
    ctx, b, _ = _gpu_build(d, n_max=n_max, bounds=bounds, algo=fmm.ALGO_CORRECTED, theta=theta)
    ctx.evaluate(b)  # the first evaluation builds the lists
    ot, od, _ = _oracle_build(d, n_max, bounds)
    m2l, p2p = O.corrected_lists(ot, od["x"], od["y"], theta)
    gm2l, gp2p = ctx.export_corrected_lists()
    _assert_lists_equal(gm2l, m2l)
    _assert_lists_equal(gp2p, p2p)


def test_corrected_lists_bit_exact_full_size_plummer1m():
    d = ic.plummer(1 << 20)
    bounds, theta = (-50.0, 50.0, -50.0, 50.0), 0.4
    ctx, b, _ = _gpu_build(d, n_max=32, bounds=bounds, algo=fmm.ALGO_CORRECTED, theta=theta)
    ctx.evaluate(b)
    ot, od, _ = _oracle_build(d, 32, bounds)
    m2l, p2p = O.corrected_lists(ot, od["x"], od["y"], theta)
    gm2l, gp2p = ctx.export_corrected_lists()
    _assert_lists_equal(gm2l, m2l)
    _assert_lists_equal(gp2p, p2p)


@pytest.mark.parametrize("p,bound", [(4, 5e-4), (6, 2e-5), (8, 2e-6)])
def test_corrected_variant_matches_oracle_and_direct_sum(p, bound):
    """p = 4, 6 and 8 run the lane-per-pair M2L and the thread-per-node M2M and L2L."""
    d = ic.plummer(20000, seed=203)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    eps, theta = 0.05, 0.3
    ctx, b, _ = _gpu_build(d, p=p, eps=eps, bounds=bounds, algo=fmm.ALGO_CORRECTED, theta=theta)
    ctx.evaluate(b)
    ot, od, _ = _oracle_build(d, 32, bounds)
    m2l, p2p = O.corrected_lists(ot, od["x"], od["y"], theta)
    ax, ay, _ = O.compute_fmm_force(od["x"], od["y"], od["m"], ot, m2l, p2p, p, eps, algo=1)
    assert np.median(_rel_diff((b.ax, b.ay), (ax, ay))) <= 1e-6
    ex, ey, _ = fmm.compute_n2_force(b.x, b.y, b.mass, eps)
    rms = np.sqrt(np.mean(_rel_diff((b.ax, b.ay), (ex, ey)) ** 2))
    assert rms < bound, rms


def test_direct_sum_matches_oracle():
    d = ic.uniform(3000, seed=204)
    ax, ay, phi = fmm.compute_n2_force(d["x"], d["y"], d["m"], 0.1, want_phi=True)
    ox, oy, oph = O.n2_force(d["x"], d["y"], d["m"], 0.1, want_phi=True)
    np.testing.assert_allclose(ax, ox, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(ay, oy, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(phi, oph, rtol=1e-12)


def test_direct_sum_kats():  # galax-core/src/lib.rs:2841-2908
    ax, ay, _ = fmm.compute_n2_force([0.0, 1.0], [0.0, 0.0], [1.0, 1.0], 0.1)
    e = 1.0 / (1.0 + 0.01) ** 1.5
    assert abs(ax[0] - e) < 1e-14 and abs(ax[1] + e) < 1e-14 and ay[0] == 0.0
    ax, _, _ = fmm.compute_n2_force([0.0, 0.0], [0.0, 0.0], [1.0, 1e6], 0.1)
    assert ax[0] == 0.0 and ax[1] == 0.0


def test_sanitize_mutates_like_reference():
    d = ic.uniform(400, seed=205, hi=10.0)
    bounds = (0.0, 10.0, 0.0, 10.0)
    ctx, b, _ = _gpu_build(d, p=2, n_max=8, bounds=bounds)
    b.x[3], b.y[5], b.mass[7], b.mass[9] = np.nan, 1e101, -1.0, np.inf
    ctx.evaluate(b)
    assert b.x[3] == 0.0 and b.y[3] == 0.0 and b.x[5] == 0.0 and b.mass[7] == 1.0 and b.mass[9] == 1.0
    assert np.isfinite(b.ax).all() and np.isfinite(b.ay).all()


def test_sanitize_every_evaluation_and_timing_switch():
    """The sanitize counters alternate between evaluations: a bad value in the
    third of several evaluations is still fixed and written back; forces with
    the timing events off equal those with them on, and the times read 0."""
    d = ic.uniform(400, seed=207, hi=10.0)
    bounds = (0.0, 10.0, 0.0, 10.0)
    ctx, b, _ = _gpu_build(d, p=2, n_max=8, bounds=bounds)
    ctx.evaluate(b)
    ax0, ay0 = b.ax.copy(), b.ay.copy()
    ctx.set_timing(False)
    ctx.evaluate(b)
    np.testing.assert_array_equal(b.ax, ax0)
    np.testing.assert_array_equal(b.ay, ay0)
    assert ctx.timings().passes_ms["total"] == 0.0 and ctx.timings().p2p_interactions > 0
    for k in range(3):
        b.mass[11 + k] = -2.0  # a bad mass in each of three more evaluations
        ctx.evaluate(b)
        assert b.mass[11 + k] == 1.0, k
        assert np.isfinite(b.ax).all()
    ctx.set_timing(True)
    ctx.evaluate(b)
    assert ctx.timings().passes_ms["total"] > 0.0


@pytest.mark.parametrize("n", [1, 2, 31, 33])
def test_small_n(n):
    d = ic.uniform(n, seed=206)
    ctx, b, _ = _gpu_build(d, p=4, bounds=(0.0, 100.0, 0.0, 100.0))
    ctx.evaluate(b)
    ex, ey, _ = fmm.compute_n2_force(b.x, b.y, b.mass, 0.1)
    if n <= 32:  # one leaf: P2P is the whole direct sum
        np.testing.assert_allclose(b.ax, ex, rtol=1e-12, atol=1e-15)
    assert np.isfinite(b.ax).all()


@pytest.mark.parametrize("n", [1, 2, 31, 33, 200])
def test_small_n_corrected(n):
    """The corrected variant on trees of one or a few leaves: the device
    traversal's lists equal the oracle's and the forces are finite; with one
    leaf they are the direct sum."""
    d = ic.uniform(n, seed=208)
    bounds = (0.0, 100.0, 0.0, 100.0)
    ctx, b, _ = _gpu_build(d, p=4, n_max=8, bounds=bounds, algo=fmm.ALGO_CORRECTED, theta=0.5)
    ctx.evaluate(b)
    assert np.isfinite(b.ax).all() and np.isfinite(b.ay).all()
    ot, od, _ = _oracle_build(d, 8, bounds)
    m2l, p2p = O.corrected_lists(ot, od["x"], od["y"], 0.5)
    gm2l, gp2p = ctx.export_corrected_lists()
    _assert_lists_equal(gm2l, m2l)
    _assert_lists_equal(gp2p, p2p)
    if n <= 8:
        ex, ey, _ = fmm.compute_n2_force(b.x, b.y, b.mass, 0.1)
        np.testing.assert_allclose(b.ax, ex, rtol=1e-12, atol=1e-15)


def test_coincident_bodies_corrected():
    """Bodies stacked on a few points: zero radii, and zero distances between
    leaves on one point (never accepted). The lists equal the oracle's and the
    forces the oracle's corrected evaluation."""
    pts = np.array([[10.0, 10.0], [10.0, 10.0], [60.0, 35.0], [60.0, 35.0], [80.0, 90.0]])
    rep = np.repeat(pts, 40, axis=0)
    n = len(rep)
    d = {"x": rep[:, 0].copy(), "y": rep[:, 1].copy(), "vx": np.zeros(n), "vy": np.zeros(n),
         "m": np.linspace(1.0, 2.0, n), "ax": np.zeros(n), "ay": np.zeros(n)}
    bounds = (0.0, 100.0, 0.0, 100.0)
    ctx, b, _ = _gpu_build(d, p=4, n_max=16, bounds=bounds, algo=fmm.ALGO_CORRECTED, theta=0.5)
    ctx.evaluate(b)
    ot, od, _ = _oracle_build(d, 16, bounds)
    m2l, p2p = O.corrected_lists(ot, od["x"], od["y"], 0.5)
    gm2l, gp2p = ctx.export_corrected_lists()
    _assert_lists_equal(gm2l, m2l)
    _assert_lists_equal(gp2p, p2p)
    ax, ay, _ = O.compute_fmm_force(od["x"], od["y"], od["m"], ot, m2l, p2p, 4, 0.1, algo=1)
    assert np.all(np.isfinite(b.ax)) and np.all(np.isfinite(b.ay))
    assert np.max(_rel_diff((b.ax, b.ay), (ax, ay))) <= 1e-9


def test_kdk_matches_oracle_simulate():
    d = ic.plummer(4000, seed=207)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    eps, p = 0.05, 4
    ctx, b, _ = _gpu_build(d, p=p, eps=eps, bounds=bounds)
    ctx.evaluate(b)
    ot, od, _ = _oracle_build(d, 32, bounds)
    m2l, p2p = O.m2l_lists(ot), O.p2p_lists(ot)
    ax, ay, _ = O.compute_fmm_force(od["x"], od["y"], od["m"], ot, m2l, p2p, p, eps)
    od["ax"][:], od["ay"][:] = ax, ay
    dt = 0.5 / np.sqrt(np.max(np.hypot(b.ax, b.ay)))
    diags = ctx.simulate(b, dt, 10, energy_every=5)
    O.simulate(od, ot, m2l, p2p, p, eps, dt, 10)
    rd = np.hypot(b.x - od["x"], b.y - od["y"]) / np.maximum(np.hypot(od["x"], od["y"]), 1e-12)
    assert np.median(rd) < 1e-9
    assert len(diags) == 2
    dg = O.diagnostics(od["x"], od["y"], od["vx"], od["vy"], od["m"], od["ax"], od["ay"])
    assert diags[-1]["kinetic"] == pytest.approx(dg["kinetic"], rel=1e-6)
    assert diags[-1]["potential_heuristic"] == pytest.approx(dg["potential"], rel=1e-6)
    # momentum and angular momentum (lib.rs:1647-1669): sums of terms of both
    # signs, so compared on the scale of the sum of their magnitudes
    mv = float(np.sum(od["m"] * np.hypot(od["vx"], od["vy"])))
    lz = float(np.sum(od["m"] * np.hypot(od["x"], od["y"]) * np.hypot(od["vx"], od["vy"])))
    assert abs(diags[-1]["momentum_x"] - dg["momentum_x"]) <= 1e-9 * mv
    assert abs(diags[-1]["momentum_y"] - dg["momentum_y"]) <= 1e-9 * mv
    assert abs(diags[-1]["angular_momentum"] - dg["angular_momentum"]) <= 1e-9 * lz
    # the potential energy 1/2 sum m_i phi_i from the FMM potentials: the
    # oracle's potentials of the final state (same tree, same algorithm)
    _, _, ophi = O.compute_fmm_force(od["x"], od["y"], od["m"], ot, m2l, p2p, p, eps, want_phi=True)
    assert diags[-1]["potential"] == pytest.approx(0.5 * float(np.sum(od["m"] * ophi)), rel=1e-9)


def test_kdk_energy_from_potentials_matches_direct_sum():
    """Device-resident KDK with the corrected variant (p = 8, theta = 0.3):
    every diagnostic against a host computation on the returned state, the
    potential energy 1/2 sum m_i phi_i against the FP64 direct sum's
    potentials (direct_potential's sign, lib.rs:1627-1635; no self term), and
    the total energy conserved to 1e-4 over 20 steps."""
    d = ic.plummer(20_000, seed=208)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    eps = 0.05
    ctx, b, _ = _gpu_build(d, p=8, eps=eps, bounds=bounds, algo=fmm.ALGO_CORRECTED, theta=0.3)
    b.phi = np.zeros(len(b))
    ctx.evaluate(b)
    dt = 2e-3
    diags = ctx.simulate(b, dt, 20, energy_every=10)
    assert len(diags) == 2
    last = diags[-1]
    ex, ey, ephi = fmm.compute_n2_force(b.x, b.y, b.mass, eps, want_phi=True)
    pe = 0.5 * float(np.sum(b.mass * ephi))
    assert last["potential"] == pytest.approx(pe, rel=1e-6)
    assert last["kinetic"] == pytest.approx(0.5 * float(np.sum(b.mass * (b.vx ** 2 + b.vy ** 2))), rel=1e-12)
    mv = float(np.sum(b.mass * np.hypot(b.vx, b.vy)))
    assert abs(last["momentum_x"] - float(np.sum(b.mass * b.vx))) <= 1e-12 * mv
    assert abs(last["momentum_y"] - float(np.sum(b.mass * b.vy))) <= 1e-12 * mv
    lz = float(np.sum(b.mass * (b.x * b.vy - b.y * b.vx)))
    assert abs(last["angular_momentum"] - lz) <= 1e-12 * float(np.sum(b.mass * np.hypot(b.x, b.y) * np.hypot(b.vx, b.vy)))
    e0 = diags[0]["kinetic"] + diags[0]["potential"]
    e1 = last["kinetic"] + last["potential"]
    assert abs(e1 - e0) <= 1e-4 * abs(e0)


def _upload_inputs(d, n_max, bounds):
    ot, od, _ = _oracle_build(d, n_max, bounds)
    return ot.export(), O.m2l_lists(ot).csr(), O.p2p_lists(ot).csr(), od


def test_upload_tree_rejects_bad_input_before_any_change():
    """upload_tree_data (galax-gpu/src/lib.rs:392-462) with a corrupted tree or
    lists is rejected on the host with InvalidArgument, before anything is
    uploaded: the tree installed before stays in use and gives the same forces."""
    d = ic.plummer(3000, seed=209)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    tree, m2l, p2p, od = _upload_inputs(d, 32, bounds)
    n = len(od["x"])
    ctx = _ctx(n, p=4, eps=0.05, bounds=bounds)
    ctx.upload_tree_data(tree, m2l, p2p)
    b = _bodies(od)
    ctx.evaluate(b)
    ax0, ay0 = b.ax.copy(), b.ay.copy()
    nn = len(tree["mass"])
    nl = int(tree["is_leaf"].sum())

    def bad_lists(csr, k, v):
        off, dat = csr[0].copy(), csr[1].copy()
        dat[k] = v
        return off, dat

    cases = {
        "p2p entry >= n_leaves": (tree, m2l, bad_lists(p2p, 1, nl)),
        "m2l entry >= n_nodes": (tree, bad_lists(m2l, 0, nn), p2p),
        "decreasing offsets": (tree, (np.where(np.arange(nn + 1) == 1, m2l[0][2] + 1, m2l[0]).astype(np.uint64),
                                      m2l[1]), p2p),
        "body_end > max_n": (dict(tree, body_end=np.where(np.arange(nn) == nn - 1, n + 1, tree["body_end"])), m2l,
                             p2p),
        "parent out of range": (dict(tree, parent=np.where(np.arange(nn) == 0, nn + 5, tree["parent"]).astype(np.int32)),
                                m2l, p2p),
        "range differs from children": (dict(tree, body_start=np.where(np.arange(nn) == nn - 1, 1,
                                                                        tree["body_start"])), m2l, p2p),
    }
    for name, (t, ml, pl) in cases.items():
        try:
            ctx.upload_tree_data(t, ml, pl)
        except fmm.GalaxError as e:
            assert e.code == -4, (name, e)
        else:
            raise AssertionError(f"upload with {name} was accepted")
    ctx._n = n  # the wrapper forgets the body count of a failed upload; the library keeps the tree
    b2 = _bodies(od)
    ctx.evaluate(b2)
    assert np.array_equal(b2.ax, ax0) and np.array_equal(b2.ay, ay0)


def test_crosscheck_against_fp64_exact():
    """GpuContext::crosscheck (galax-gpu/src/lib.rs:623-667): the mixed form
    against galax-core's arithmetic (FP64_EXACT) on the same tree. The metric
    is the reference's; bodies.ax/ay end up holding the FP64_EXACT forces;
    a threshold below the error gives ThresholdExceeded (the reference panics)."""
    d = ic.plummer(20_000, seed=210)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    ctx, b, _ = _gpu_build(d, p=6, eps=0.05, bounds=bounds, precision=fmm.PREC_MIXED_FP32)
    ctx.evaluate(b)
    am, bm = b.ax.copy(), b.ay.copy()
    mx, mean, cnt = ctx.crosscheck(b, 0.5)
    assert cnt == len(b)
    ctx_e, be, _ = _gpu_build(d, p=6, eps=0.05, bounds=bounds, precision=fmm.PREC_FP64_EXACT)
    ctx_e.evaluate(be)
    assert np.array_equal(b.ax, be.ax) and np.array_equal(b.ay, be.ay)
    den = np.maximum(np.maximum(np.hypot(be.ax, be.ay), np.hypot(am, bm)), 1e-30)
    rel = np.hypot(am - be.ax, bm - be.ay) / den
    assert mx == pytest.approx(float(rel.max()), rel=1e-12)
    assert mean == pytest.approx(float(rel.mean()), rel=1e-9)
    assert 0.0 < mx < 1e-3
    with pytest.raises(fmm.GalaxError) as ei:
        ctx.crosscheck(b, mx / 2)
    assert ei.value.code == -7
    # the context's own form is unchanged afterwards
    ctx.evaluate(b)
    assert np.array_equal(b.ax, am) and np.array_equal(b.ay, bm)


def test_errors_are_codes_not_crashes():
    with pytest.raises(fmm.GalaxError):
        fmm.GpuContext(fmm.GpuConfig(max_n=10, p=17))
    ctx = _ctx(10)
    b = _bodies(ic.uniform(10))
    with pytest.raises(fmm.GalaxError):
        ctx.evaluate(b)  # no tree yet
    big = _bodies(ic.uniform(11))
    with pytest.raises(fmm.GalaxError):
        ctx.build_tree(big)  # above max_n


# ------------------------------------------------------ mixed precision ----

@pytest.mark.parametrize("p", [2, 6, 8])
def test_fp64_mfma_m2l_matches_exact_order(p):
    """The FP64 form (M2L grouped by key on FP64 MFMA) against FP64_EXACT
    (galax-core's per-pair order): the same operators summed in another order."""
    d = ic.plummer(50_000, seed=303)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    out = {}
    for prec in (fmm.PREC_FP64, fmm.PREC_FP64_EXACT):
        ctx, b, _ = _gpu_build(d, p=p, eps=0.02, bounds=bounds, precision=prec)
        ctx.evaluate(b)
        out[prec] = (b.ax.copy(), b.ay.copy(), b.phi.copy())
    a, e = out[fmm.PREC_FP64], out[fmm.PREC_FP64_EXACT]
    rd = _rel_diff(a[:2], e[:2])
    assert rd.max() <= 1e-10, rd.max()
    np.testing.assert_allclose(a[2], e[2], rtol=1e-11, atol=0.0)

@pytest.mark.parametrize("p", [12, 16])
def test_high_orders_all_forms(p):
    """p = 12 and 16 (the largest): the FP64 form (and at p = 12 the mixed
    forms) against FP64_EXACT, and FP64_EXACT against the oracle, on a small
    Plummer sphere. These orders run the register-heavy variants of every
    kernel. (At p = 16 the scaled FP32 operators reach (2p)! ~ 3e35, at the
    edge of FP32's range, so the mixed forms are checked at p = 12 only. There
    the reference algorithm's key cache makes the high-order M2L terms large
    and cancelling, so FP32 loses more than at p <= 8: measured 1.9e-5 RMS vs
    FP64; the bound is 1e-4.)"""
    d = ic.plummer(3000, seed=306)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    eps = 0.02
    out = {}
    forms = (fmm.PREC_FP64_EXACT, fmm.PREC_FP64) + ((fmm.PREC_MIXED_FP32, fmm.PREC_MIXED_BF16X3) if p <= 12 else ())
    for prec in forms:
        ctx, b, _ = _gpu_build(d, p=p, eps=eps, bounds=bounds, precision=prec)
        ctx.evaluate(b)
        assert np.isfinite(b.ax).all() and np.isfinite(b.ay).all(), prec
        out[prec] = (b.ax.copy(), b.ay.copy())
    e = out[fmm.PREC_FP64_EXACT]
    assert _rel_diff(out[fmm.PREC_FP64], e).max() <= 1e-10
    for prec in forms[2:]:
        rd = _rel_diff(out[prec], e)
        assert np.sqrt(np.mean(rd ** 2)) <= 1e-4, (prec, np.sqrt(np.mean(rd ** 2)))
    ot, od, _ = _oracle_build(d, 32, bounds)
    ax, ay, _ = O.compute_fmm_force(od["x"], od["y"], od["m"], ot, O.m2l_lists(ot), O.p2p_lists(ot), p, eps,
                                    want_phi=True)
    rd = _rel_diff(e, (ax, ay))
    assert np.median(rd) <= 1e-6 and rd.max() <= 1e-9, (np.median(rd), rd.max())


def _forms_run(d, p, eps, bounds, algo=fmm.ALGO_REFERENCE, n_max=32):
    out = {}
    for prec in (fmm.PREC_FP64, fmm.PREC_MIXED_FP32, fmm.PREC_MIXED_BF16X3):
        ctx, b, _ = _gpu_build(d, p=p, eps=eps, bounds=bounds, precision=prec, algo=algo, n_max=n_max, theta=0.3)
        ctx.evaluate(b)
        out[prec] = (b.ax.copy(), b.ay.copy(), b.phi.copy())
    return out


@pytest.mark.parametrize("p", [4, 6, 8])
def test_mixed_forms_match_fp64_reference_algo(p):
    """FP32-MFMA M2L + FP32 P2P, and the BF16x3 M2L, against the FP64 path of the
    same algorithm: RMS relative force difference <= 1e-5 (the mixed forms'
    accuracy target)."""
    d = ic.plummer(50_000, seed=301)
    res = _forms_run(d, p, 0.02, (-50.0, 50.0, -50.0, 50.0))
    ref = res[fmm.PREC_FP64]
    for prec in (fmm.PREC_MIXED_FP32, fmm.PREC_MIXED_BF16X3):
        rd = _rel_diff(res[prec][:2], ref[:2])
        assert np.sqrt(np.mean(rd ** 2)) <= 1e-5, (prec, np.sqrt(np.mean(rd ** 2)), rd.max())
        # potentials: RMS relative difference (a few bodies whose far-field
        # potential dominates see the FP32 M2L rounding at ~1e-4)
        pr = np.abs(res[prec][2] - ref[2]) / np.abs(ref[2])
        assert np.sqrt(np.mean(pr ** 2)) <= 1e-5, (prec, np.sqrt(np.mean(pr ** 2)), pr.max())


SYM_CASES = [
    ("plummer_nmax32_ref", lambda: ic.plummer(40_000, seed=311), 32, fmm.ALGO_REFERENCE),
    ("plummer_nmax8_ref", lambda: ic.plummer(12_000, seed=312), 8, fmm.ALGO_REFERENCE),
    ("uniform_nmax13_ref", lambda: ic.uniform(9_001, seed=313), 13, fmm.ALGO_REFERENCE),
    ("plummer_nmax32_corrected", lambda: ic.plummer(30_000, seed=314), 32, fmm.ALGO_CORRECTED),
    ("galaxy_nmax16_corrected", lambda: ic.galaxy(8_192, seed=315), 16, fmm.ALGO_CORRECTED),
    # leaves of more than 32 bodies: the pairs of their halves
    ("plummer_nmax64_ref", lambda: ic.plummer(40_000, seed=316), 64, fmm.ALGO_REFERENCE),
    ("uniform_nmax48_ref", lambda: ic.uniform(9_001, seed=317), 48, fmm.ALGO_REFERENCE),
    ("plummer_nmax64_corrected", lambda: ic.plummer(30_000, seed=318), 64, fmm.ALGO_CORRECTED),
    ("galaxy_nmax33_corrected", lambda: ic.galaxy(8_192, seed=319), 33, fmm.ALGO_CORRECTED),
]


@pytest.mark.parametrize("name,gen,n_max,algo", SYM_CASES, ids=[c[0] for c in SYM_CASES])
def test_symmetric_near_field_matches_gather_kernel(name, gen, n_max, algo, monkeypatch):
    """The FP32 forms' near field evaluates each unordered leaf pair once
    (k_p2p_sym, the reference's pair rule a < b, galax-core/src/lib.rs:1415-1437;
    leaves of more than 32 bodies as the pairs of their halves); the gather
    kernel (GALAX_NEAR_GATHER=1) evaluates every directed pair. Both
    are FP32 on leaf-relative coordinates: forces and potentials agree to FP32
    rounding, on full and partly filled leaves, in both variants. Both stay
    within 1e-5 RMS of FP64."""
    d = gen()
    bounds = (-50.0, 50.0, -50.0, 50.0) if name.startswith(("plummer", "galaxy")) else (0.0, 100.0, 0.0, 100.0)
    out = {}
    for label, env, prec in (("sym", "0", fmm.PREC_MIXED_FP32), ("gather", "1", fmm.PREC_MIXED_FP32),
                             ("fp64", "0", fmm.PREC_FP64)):
        monkeypatch.setenv("GALAX_NEAR_GATHER", env)
        ctx, b, _ = _gpu_build(d, p=6, eps=0.05, bounds=bounds, precision=prec, algo=algo, n_max=n_max, theta=0.4)
        ctx.evaluate(b)
        assert np.isfinite(b.ax).all() and np.isfinite(b.phi).all(), label
        out[label] = (b.ax.copy(), b.ay.copy(), b.phi.copy())
    rs = _rel_diff(out["sym"][:2], out["gather"][:2])
    assert np.sqrt(np.mean(rs ** 2)) <= 2e-6 and np.median(rs) <= 1e-6, (np.sqrt(np.mean(rs ** 2)), rs.max())
    ps = np.abs(out["sym"][2] - out["gather"][2]) / np.abs(out["gather"][2])
    assert np.sqrt(np.mean(ps ** 2)) <= 2e-6, np.sqrt(np.mean(ps ** 2))
    for label in ("sym", "gather"):
        r = _rel_diff(out[label][:2], out["fp64"][:2])
        assert np.sqrt(np.mean(r ** 2)) <= 1e-5, (label, np.sqrt(np.mean(r ** 2)))


@pytest.mark.parametrize("algo,n_max", [(fmm.ALGO_REFERENCE, 32), (fmm.ALGO_CORRECTED, 32), (fmm.ALGO_REFERENCE, 64)],
                         ids=["reference", "corrected", "reference_nmax64"])
def test_symmetric_near_field_after_rebuild_with_potential(algo, n_max):
    """A handle that evaluated forces and potentials on one tree, then builds a
    larger one (more leaf pairs, more work items) and evaluates again: its
    results equal a fresh handle's bit for bit. (The potentials' near-field
    slots were once sized at their first use only, so the second tree's
    evaluation wrote past them.)"""
    bounds = (-50.0, 50.0, -50.0, 50.0)
    small, big = ic.plummer(12_000, seed=321), ic.plummer(40_000, seed=322)
    kw = dict(p=6, eps=0.05, n_max=n_max, bounds=bounds, precision=fmm.PREC_MIXED_FP32, algo=algo, theta=0.4)
    ctx = _ctx(len(big["x"]), **kw)
    b = _bodies(small)
    ctx.build_tree(b)
    b.phi = np.zeros(len(b))
    ctx.evaluate(b)
    b2 = _bodies(big)
    ctx.build_tree(b2)
    b2.phi = np.zeros(len(b2))
    ctx.evaluate(b2)
    fresh, b3, _ = _gpu_build(big, **kw)
    b3.phi = np.zeros(len(b3))
    fresh.evaluate(b3)
    for u, v in ((b2.ax, b3.ax), (b2.ay, b3.ay), (b2.phi, b3.phi)):
        assert np.isfinite(u).all()
        assert np.array_equal(u, v)


def test_mixed_forms_match_fp64_plummer1m_p6():
    """The full 2^20-body Plummer workload at p = 6 (size-independent check:
    mixed vs FP64 of the same algorithm)."""
    d = ic.plummer(1 << 20)
    eps = 0.05
    res = _forms_run(d, 6, eps, (-50.0, 50.0, -50.0, 50.0))
    ref = res[fmm.PREC_FP64]
    assert np.isfinite(ref[0]).all()
    for prec in (fmm.PREC_MIXED_FP32, fmm.PREC_MIXED_BF16X3):
        rd = _rel_diff(res[prec][:2], ref[:2])
        assert np.sqrt(np.mean(rd ** 2)) <= 1e-5, (prec, np.sqrt(np.mean(rd ** 2)))


def test_corrected_plummer1m_within_1e5_of_direct_sum():
    """The full 2^20-body Plummer workload as the bench runs it (galax-cli's
    softening of the sorted bodies) with the corrected variant at p = 8,
    theta = 0.35, FP64 and MIXED_FP32, against the FP64 direct sum on sampled
    targets (size-independent: the direct sum is exact for each of them). The
    RMS of the relative error is dominated by the few bodies whose force
    nearly cancels, so it varies between samples: it must stay within 1e-5 on
    each of three independent 10,000-draw samples (measured 1.1e-6 .. 5.2e-6;
    at theta = 0.4 one of them gives 1.6e-5). The median, the 99th
    percentile and RMS |da| / RMS |a| are checked as well."""
    d = ic.plummer(1 << 20)
    bounds = (-50.0, 50.0, -50.0, 50.0)
    out = {}
    for prec in (fmm.PREC_FP64, fmm.PREC_MIXED_FP32):
        ctx, b, _ = _gpu_build(d, p=8, bounds=bounds, algo=fmm.ALGO_CORRECTED, theta=0.35, precision=prec)
        eps = fmm.default_softening(b.x, b.y)  # galax-cli/src/main.rs:182-190
        ctx.set_softening(eps)
        ctx.evaluate(b)
        out[prec] = (b.ax.copy(), b.ay.copy())
    for seed in (4242, 77, 2026):
        si = np.unique((ic.SplitMix64(seed).next_u64(10_000) % np.uint64(len(b))).astype(np.int64))
        ex, ey, _ = fmm.compute_n2_force(b.x, b.y, b.mass, eps, targets=si.astype(np.uint64))
        for prec, (ax, ay) in out.items():
            rd = _rel_diff((ax[si], ay[si]), (ex, ey))
            print(f"corrected p=8 theta=0.35 prec={prec} seed={seed}: rms {np.sqrt(np.mean(rd ** 2)):.3e}", flush=True)
            assert np.sqrt(np.mean(rd ** 2)) <= 1e-5, (prec, seed, np.sqrt(np.mean(rd ** 2)))
            dabs = np.hypot(ax[si] - ex, ay[si] - ey)
            norm = np.sqrt(np.mean(dabs ** 2)) / np.sqrt(np.mean(ex ** 2 + ey ** 2))
            assert norm <= 1e-5 and np.median(rd) <= 1e-6 and np.percentile(rd, 99) <= 1e-4, (
                prec, seed, norm, np.median(rd), np.percentile(rd, 99))


def test_mixed_forms_corrected_algo():
    d = ic.plummer(30_000, seed=302)
    res = _forms_run(d, 6, 0.05, (-50.0, 50.0, -50.0, 50.0), algo=fmm.ALGO_CORRECTED)
    rd = _rel_diff(res[fmm.PREC_MIXED_FP32][:2], res[fmm.PREC_FP64][:2])
    assert np.sqrt(np.mean(rd ** 2)) <= 1e-5
