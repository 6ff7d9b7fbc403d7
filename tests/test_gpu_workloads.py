"""The exponential-disk (2^22 bodies), galaxy (2^18) and clustered (2^24,
n_max = 64) workloads at their full sizes on the GPU, through the C ABI.

Checks that do not need an O(N^2) oracle at these sizes:
- the mixed forms against the FP64 form of the same algorithm (RMS relative
  force difference <= 1e-5, the mixed forms' accuracy target);
- the FP64 direct sum on sampled targets (exact for each sampled target);
- central differences of the forward map (the VJP), the adjoint of the
  direct sum at N = 8192 (oracle), and checkpointing invariance of a
  leapfrog gradient;
- the tree and body permutation bit for bit against the oracle's CPU build
  (Tree::build, galax-core/src/lib.rs:163-361) at 2^24 bodies.
"""
import numpy as np
import pytest

from galax_amd import fmm, ic
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP64, MIXED, BF16X3 = fmm.PREC_FP64, fmm.PREC_MIXED_FP32, fmm.PREC_MIXED_BF16X3


def _samples(n, seed, k=10_000):
    return np.unique((ic.SplitMix64(seed).next_u64(k) % np.uint64(n)).astype(np.int64))


def _rel(a, e):
    return np.hypot(a[0] - e[0], a[1] - e[1]) / np.maximum(np.hypot(e[0], e[1]), 1e-300)


def _rms(v):
    return float(np.sqrt(np.mean(v ** 2)))


def _build(d, spec, p, prec, algo=fmm.ALGO_REFERENCE, theta=0.3, eps=None):
    """A context with the workload's n_max and bounds, its tree built from d,
    and galax-cli's softening of the sorted bodies (galax-cli/src/main.rs:182-190)
    unless eps is given."""
    n = len(d["x"])
    bounds = spec["bounds"]
    if bounds == "data":
        bounds = (float(d["x"].min()), float(d["x"].max()), float(d["y"].min()), float(d["y"].max()))
    ctx = fmm.GpuContext(fmm.GpuConfig(max_n=n, p=p, eps=0.1, n_max=spec["n_max"], bounds=bounds, algo=algo,
                                       theta=theta, precision=prec))
    b = fmm.BodiesSoA.from_dict(d)
    perm = ctx.build_tree(b)
    ctx.set_softening(fmm.default_softening(b.x, b.y) if eps is None else eps)
    return ctx, b, perm


# ----------------------------------------------------------- disk4m ----

@pytest.fixture(scope="module")
def disk4m():
    d, spec = ic.make("disk4m")
    return d, spec


@pytest.mark.parametrize("p", [2, 4, 6, 8])
def test_disk4m_precision_sweep(disk4m, p):
    """Exponential disk, 2^22 bodies, n_max = 32, reference algorithm: FP32-MFMA
    M2L + FP32 P2P and the BF16x3 M2L against FP64 (RMS <= 1e-5). Against the
    FP64 direct sum on 10,000 sampled targets every form carries the reference
    algorithm's own error (its key cache and id-ordered L2L; measured RMS
    ~1.1 at p <= 6, ~36 at p = 8), the same in every form to 1e-4 relative."""
    d, spec = disk4m
    print(f"disk4m p={p}", flush=True)
    out = {}
    for prec in (FP64, MIXED, BF16X3):
        ctx, b, _ = _build(d, spec, p, prec)
        ctx.evaluate(b)
        assert np.isfinite(b.ax).all() and np.isfinite(b.ay).all(), prec
        out[prec] = (b.ax.copy(), b.ay.copy())
        ctx.close()
    ref = out[FP64]
    for prec in (MIXED, BF16X3):
        r = _rms(_rel(out[prec], ref))
        assert r <= 1e-5, (prec, r)
    si = _samples(len(b), 4242)
    eps = fmm.default_softening(b.x, b.y)
    ex, ey, _ = fmm.compute_n2_force(b.x, b.y, b.mass, eps, targets=si.astype(np.uint64))
    e64 = _rms(_rel((ref[0][si], ref[1][si]), (ex, ey)))
    assert np.isfinite(e64) and e64 > 0
    for prec in (MIXED, BF16X3):
        e = _rms(_rel((out[prec][0][si], out[prec][1][si]), (ex, ey)))
        assert abs(e - e64) <= 1e-4 * e64, (prec, e, e64)


def test_disk4m_corrected_within_1e5_of_direct_sum(disk4m):
    """The corrected variant (p = 6, theta = 0.3), FP64 and MIXED_FP32, against
    the FP64 direct sum: RMS relative error <= 1e-5 on three independent
    10,000-target samples."""
    d, spec = disk4m
    out = {}
    for prec in (FP64, MIXED):
        ctx, b, _ = _build(d, spec, 6, prec, algo=fmm.ALGO_CORRECTED, theta=0.3)
        ctx.evaluate(b)
        out[prec] = (b.ax.copy(), b.ay.copy())
        ctx.close()
    eps = fmm.default_softening(b.x, b.y)
    for seed in (4242, 77, 2026):
        si = _samples(len(b), seed)
        ex, ey, _ = fmm.compute_n2_force(b.x, b.y, b.mass, eps, targets=si.astype(np.uint64))
        for prec, (ax, ay) in out.items():
            r = _rms(_rel((ax[si], ay[si]), (ex, ey)))
            print(f"disk4m corrected p=6 prec={prec} seed={seed}: rms rel vs direct {r:.3e}", flush=True)
            assert r <= 1e-5, (prec, seed, r)


# -------------------------------------------------------- galaxy262k ----

@pytest.fixture(scope="module")
def galaxy262k():
    return ic.make("galaxy262k")


def _cotangents(n, seed):
    g = ic.SplitMix64(seed)
    return g.normal(n), g.normal(n)


def _forward(ctx, b, x, y, m):
    c = fmm.BodiesSoA(x.copy(), y.copy(), b.vx, b.vy, m.copy(), np.zeros(len(b)), np.zeros(len(b)), None)
    ctx.evaluate(c)
    return c.ax, c.ay


@pytest.mark.parametrize("algo", [fmm.ALGO_REFERENCE, fmm.ALGO_CORRECTED], ids=["reference", "corrected"])
def test_galaxy262k_vjp_central_differences(galaxy262k, algo):
    """VJP of the accelerations on the galaxy preset (p = 6, FP64, frozen tree)
    against central differences with h = 1e-6 on 64 random coordinates of
    x, y and m (measured worst relative error: 4.5e-8 reference, 2.1e-7
    corrected)."""
    d, spec = galaxy262k
    ctx, b, _ = _build(d, spec, 6, FP64, algo=algo)
    _forward(ctx, b, b.x, b.y, b.mass)  # the corrected lists are frozen at the first evaluation
    n = len(b)
    axb, ayb = _cotangents(n, 11)
    gx, gy, gm = ctx.vjp(b, axb, ayb)
    rr = ic.SplitMix64(12)
    idx = (rr.next_u64(64) % np.uint64(n)).astype(np.int64)
    which = (rr.next_u64(64) % np.uint64(3)).astype(np.int64)
    h = 1e-6
    worst = 0.0
    floor = 1e-3 * max(np.abs(gx).max(), np.abs(gy).max())
    for k, w in zip(idx, which):
        vals = []
        for sgn in (1.0, -1.0):
            arrs = [b.x.copy(), b.y.copy(), b.mass.copy()]
            arrs[w][k] += sgn * h
            fx, fy = _forward(ctx, b, *arrs)
            vals.append(np.dot(fx, axb) + np.dot(fy, ayb))
        fd = (vals[0] - vals[1]) / (2 * h)
        g = (gx, gy, gm)[w][k]
        worst = max(worst, abs(fd - g) / max(abs(fd), abs(g), floor))
    print(f"galaxy262k FD worst relative error (algo {algo}): {worst:.3e}", flush=True)
    assert worst < 1e-5, worst


@pytest.mark.parametrize("algo", [fmm.ALGO_REFERENCE, fmm.ALGO_CORRECTED], ids=["reference", "corrected"])
def test_galaxy262k_mixed_vjp_close_to_fp64(galaxy262k, algo):
    """The VJP with the FP32 forward (MIXED_FP32: FP32 near-field adjoint on
    leaf-relative coordinates) against the FP64 form's, in norm."""
    d, spec = galaxy262k
    axb, ayb = _cotangents(len(d["x"]), 13)
    out = {}
    for prec in (FP64, MIXED):
        ctx, b, _ = _build(d, spec, 6, prec, algo=algo)
        _forward(ctx, b, b.x, b.y, b.mass)
        out[prec] = ctx.vjp(b, axb, ayb)
        ctx.close()
    for g, g64 in zip(out[MIXED], out[FP64]):
        assert np.isfinite(g).all()
        rel = np.linalg.norm(g - g64) / np.linalg.norm(g64)
        print(f"galaxy262k mixed vs fp64 VJP (algo {algo}): {rel:.3e}", flush=True)
        assert rel <= 1e-5, rel


def test_galaxy_preset_direct_sum_adjoint_8192():
    """The galaxy preset at N = 8192 (a 1000-mass centre and 1e-3 test
    bodies): the corrected variant's VJP (p = 6, theta = 0.3) approximates the
    closed-form adjoint of the FP64 direct sum (oracle)."""
    d, spec = ic.make("galaxy262k", n=8192)
    ctx, b, _ = _build(d, spec, 6, FP64, algo=fmm.ALGO_CORRECTED)
    eps = fmm.default_softening(b.x, b.y)
    _forward(ctx, b, b.x, b.y, b.mass)
    axb, ayb = _cotangents(len(b), 14)
    gx, gy, gm = ctx.vjp(b, axb, ayb)
    ox, oy, om = O.n2_vjp(b.x, b.y, b.mass, eps, axb, ayb, threads=8)
    for name, g, o in (("x", gx, ox), ("y", gy, oy), ("m", gm, om)):
        rel = np.linalg.norm(g - o) / np.linalg.norm(o)
        print(f"galaxy8192 VJP vs direct-sum adjoint ({name}): {rel:.3e}", flush=True)
        assert rel < 1e-3, (name, rel)


def test_galaxy262k_checkpointed_leapfrog50():
    """A 50-step leapfrog on the galaxy preset (corrected variant, p = 6, dt =
    1e-4), loss = MSE to a target perturbed by N(0, 0.01): the gradient with
    checkpoints every 10 steps equals the one without checkpoints (the
    recomputed segments run the same deterministic kernels), in the FP32
    forward form and in FP64; in FP64 it matches central differences at two
    coordinates."""
    torch = pytest.importorskip("torch")
    from galax_amd import autograd as GA

    d, spec = ic.make("galaxy262k")
    n = len(d["x"])
    dev = torch.device("cuda:0")
    dt, steps = 1e-4, 50
    for prec in (MIXED, FP64):
        ctx, b, _ = _build(d, spec, 6, prec, algo=fmm.ALGO_CORRECTED)
        x = torch.tensor(b.x, device=dev)
        y = torch.tensor(b.y, device=dev)
        m = torch.tensor(b.mass, device=dev)
        vx = torch.tensor(b.vx, device=dev)
        vy = torch.tensor(b.vy, device=dev)
        tg = ic.SplitMix64(13)
        tx = x + 0.01 * torch.tensor(tg.normal(n), device=dev)
        ty = y + 0.01 * torch.tensor(tg.normal(n), device=dev)

        def loss_of(x0, every, with_grad=True):
            with torch.set_grad_enabled(with_grad):
                xf, yf, _, _ = GA.leapfrog(x0, y, vx, vy, m, ctx, dt, steps, checkpoint_every=every)
                return ((xf - tx) ** 2 + (yf - ty) ** 2).mean()

        grads = []
        for every in (10, 0):
            x0 = x.clone().requires_grad_(True)
            (g,) = torch.autograd.grad(loss_of(x0, every), (x0,))
            grads.append(g)
        assert torch.isfinite(grads[0]).all()
        assert torch.equal(grads[0], grads[1]), prec
        print(f"galaxy262k leapfrog50 prec={prec}: |grad| = {float(grads[0].norm()):.4e}", flush=True)
        if prec == FP64:
            h = 1e-6
            for k in (5, 200_000):
                xp, xm = x.clone(), x.clone()
                xp[k] += h
                xm[k] -= h
                fd = (loss_of(xp, 10, False) - loss_of(xm, 10, False)).item() / (2 * h)
                assert fd == pytest.approx(grads[0][k].item(), rel=1e-4, abs=1e-12), (k, fd, grads[0][k].item())
        ctx.close()


# ------------------------------------------------------ clustered16m ----

def test_clustered16m_tree_bit_exact_and_forms():
    """64 Plummer clusters, 2^24 bodies, n_max = 64, bounds = data bbox, p = 8:
    the device tree and permutation equal the oracle's CPU Tree::build bit for
    bit; MIXED_FP32 against FP64 RMS <= 1e-5; the mixed form's VJP against
    the FP64 form's (a Gaussian cotangent) within 1e-5 in norm."""
    d, spec = ic.make("clustered16m")
    n = len(d["x"])
    print(f"clustered16m n={n}", flush=True)
    out, vjps = {}, {}
    axb, ayb = _cotangents(n, 21)
    for prec in (FP64, MIXED):
        ctx, b, perm = _build(d, spec, 8, prec)
        if prec == FP64:
            bounds = spec["bounds"]  # the data bbox (ic.make resolves "data")
            od = {k: np.array(d[k], np.float64) for k in ("x", "y", "vx", "vy", "m", "ax", "ay")}
            ot, operm = O.tree_build(od, 64, bounds, return_perm=True)
            assert np.array_equal(perm, operm)
            assert np.array_equal(b.x.view(np.uint64), od["x"].view(np.uint64))
            gt, _, _ = ctx.export_tree()
            want = ot.export()
            for k, v in want.items():
                a = np.asarray(gt[k])
                assert a.shape == v.shape, k
                if v.dtype == np.float64:
                    assert np.array_equal(a.view(np.uint64), v.view(np.uint64)), k
                else:
                    assert np.array_equal(a.astype(np.int64), v.astype(np.int64)), k
            del ot, od, want, gt
        ctx.evaluate(b)
        assert np.isfinite(b.ax).all() and np.isfinite(b.ay).all(), prec
        out[prec] = (b.ax.copy(), b.ay.copy())
        vjps[prec] = ctx.vjp(b, axb, ayb)
        ctx.close()
    r = _rms(_rel(out[MIXED], out[FP64]))
    print(f"clustered16m p=8 mixed vs fp64 rms {r:.3e}", flush=True)
    assert r <= 1e-5, r
    for g, g64 in zip(vjps[MIXED], vjps[FP64]):
        rel = np.linalg.norm(g - g64) / np.linalg.norm(g64)
        print(f"clustered16m VJP mixed vs fp64: {rel:.3e}", flush=True)
        assert np.isfinite(g).all() and rel <= 1e-5, rel
