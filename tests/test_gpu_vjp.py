"""The VJP of the FMM accelerations (galax_fmm_vjp) on the GPU: central finite
differences in FP64 (h = 1e-6) of the device forward on a frozen tree, the
closed-form adjoint of the direct sum (oracle) where the FMM is exact, and the
torch.autograd wrapper with a checkpointed leapfrog."""
import numpy as np
import pytest

from galax_amd import fmm, ic
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _ctx(d, p, eps, algo, n_max=32, bounds=(-50.0, 50.0, -50.0, 50.0), theta=0.3):
    n = len(d["x"])
    ctx = fmm.GpuContext(fmm.GpuConfig(max_n=n, p=p, eps=eps, n_max=n_max, bounds=bounds, algo=algo, theta=theta))
    b = fmm.BodiesSoA.from_dict(d)
    ctx.build_tree(b)
    return ctx, b


def _forward(ctx, b, x=None, y=None, m=None):
    c = fmm.BodiesSoA((x if x is not None else b.x).copy(), (y if y is not None else b.y).copy(), b.vx, b.vy,
                      (m if m is not None else b.mass).copy(), np.zeros(len(b)), np.zeros(len(b)), None)
    ctx.evaluate(c)
    return c.ax, c.ay


@pytest.mark.parametrize("algo,p", [(fmm.ALGO_REFERENCE, 6), (fmm.ALGO_CORRECTED, 6), (fmm.ALGO_CORRECTED, 8)])
def test_vjp_matches_central_differences(algo, p):
    d = ic.plummer(4096, seed=401)
    eps = 0.05
    ctx, b = _ctx(d, p, eps, algo)
    _forward(ctx, b)  # corrected lists are frozen at the first evaluation
    r = ic.SplitMix64(9)
    axb, ayb = r.normal(len(b)), r.normal(len(b))
    xb, yb, mb = ctx.vjp(b, axb, ayb)
    h = 1e-6
    idx = (r.next_u64(22) % np.uint64(len(b))).astype(np.int64)
    worst = 0.0
    for k in idx:
        for which, g in (("x", xb), ("y", yb), ("m", mb)):
            arrs = {"x": b.x.copy(), "y": b.y.copy(), "m": b.mass.copy()}
            arrs[which][k] += h
            fp = _forward(ctx, b, **arrs)
            arrs[which][k] -= 2 * h
            fm = _forward(ctx, b, **arrs)
            fd = (np.dot(fp[0] - fm[0], axb) + np.dot(fp[1] - fm[1], ayb)) / (2 * h)
            scale = max(abs(fd), abs(g[k]), 1e-3 * np.abs(g).max())
            worst = max(worst, abs(fd - g[k]) / scale)
    assert worst < 1e-5, worst


SYM_VJP_CASES = [
    ("plummer_nmax32_ref", lambda: ic.plummer(12_000, seed=411), 32, fmm.ALGO_REFERENCE),
    ("plummer_nmax64_ref", lambda: ic.plummer(20_000, seed=412), 64, fmm.ALGO_REFERENCE),
    ("uniform_nmax13_ref", lambda: ic.uniform(9_001, seed=413), 13, fmm.ALGO_REFERENCE),
    ("plummer_nmax32_corrected", lambda: ic.plummer(10_000, seed=414), 32, fmm.ALGO_CORRECTED),
    ("galaxy_nmax48_corrected", lambda: ic.galaxy(8_192, seed=415), 48, fmm.ALGO_CORRECTED),
]


@pytest.mark.parametrize("name,gen,n_max,algo", SYM_VJP_CASES, ids=[c[0] for c in SYM_VJP_CASES])
def test_symmetric_near_adjoint_matches_gather_adjoint(name, gen, n_max, algo, monkeypatch):
    """The FP32 forms' near-field adjoint evaluates each unordered leaf pair
    once (k_vjp_sym, the symmetric near field's items; leaves of more than 32
    bodies as their halves); with GALAX_NEAR_GATHER=1 it is the gather kernel,
    every directed pair. Both are FP32 on leaf-relative coordinates: the
    cotangents agree to FP32 rounding (relative norm 2e-6 on these cases,
    1e-5 where a 1000:1e-3 mass ratio makes mbar cancel), and both stay within
    1e-5 of FP64."""
    d = gen()
    bounds = (0.0, 100.0, 0.0, 100.0) if name.startswith("uniform") else (-50.0, 50.0, -50.0, 50.0)
    r = ic.SplitMix64(17)
    axb, ayb = r.normal(len(d["x"])), r.normal(len(d["x"]))
    out = {}
    for label, env, prec in (("sym", "0", fmm.PREC_MIXED_FP32), ("gather", "1", fmm.PREC_MIXED_FP32),
                             ("fp64", "0", fmm.PREC_FP64)):
        monkeypatch.setenv("GALAX_NEAR_GATHER", env)
        n = len(d["x"])
        ctx = fmm.GpuContext(fmm.GpuConfig(max_n=n, p=6, eps=0.05, n_max=n_max, bounds=bounds, algo=algo,
                                           theta=0.4, precision=prec))
        b = fmm.BodiesSoA.from_dict(d)
        ctx.build_tree(b)
        _forward(ctx, b)
        g = ctx.vjp(b, axb, ayb)
        assert all(np.isfinite(v).all() for v in g), label
        out[label] = g
    tol = 1e-5 if name.startswith("galaxy") else 2e-6
    for k in range(3):
        ref = np.linalg.norm(out["gather"][k])
        print(name, k, "sym-gather", np.linalg.norm(out["sym"][k] - out["gather"][k]) / ref,
              "sym-fp64", np.linalg.norm(out["sym"][k] - out["fp64"][k]) / np.linalg.norm(out["fp64"][k]),
              "gather-fp64", np.linalg.norm(out["gather"][k] - out["fp64"][k]) / np.linalg.norm(out["fp64"][k]),
              flush=True)
        assert np.linalg.norm(out["sym"][k] - out["gather"][k]) <= tol * ref, (k, name)
        for label in ("sym", "gather"):
            assert np.linalg.norm(out[label][k] - out["fp64"][k]) <= 1e-5 * np.linalg.norm(out["fp64"][k]), (label, k)


@pytest.mark.parametrize("algo", [fmm.ALGO_REFERENCE, fmm.ALGO_CORRECTED])
def test_mixed_form_vjp_matches_fp64_vjp(algo):
    """The mixed forms run the near-field adjoint in FP32 on leaf-relative
    coordinates (their forward near field's precision) and the rest in FP64:
    the cotangents must match the FP64 form's to 1e-5 in norm."""
    d = ic.plummer(8192, seed=405)
    r = ic.SplitMix64(11)
    axb, ayb = r.normal(len(d["x"])), r.normal(len(d["x"]))
    out = {}
    for prec in (fmm.PREC_FP64, fmm.PREC_MIXED_FP32, fmm.PREC_MIXED_BF16X3):
        n = len(d["x"])
        ctx = fmm.GpuContext(fmm.GpuConfig(max_n=n, p=6, eps=0.05, n_max=32, bounds=(-50.0, 50.0, -50.0, 50.0),
                                           algo=algo, theta=0.3, precision=prec))
        b = fmm.BodiesSoA.from_dict(d)
        ctx.build_tree(b)
        _forward(ctx, b)
        out[prec] = ctx.vjp(b, axb, ayb)
    for prec in (fmm.PREC_MIXED_FP32, fmm.PREC_MIXED_BF16X3):
        for g, g64 in zip(out[prec], out[fmm.PREC_FP64]):
            rel = np.linalg.norm(g - g64) / np.linalg.norm(g64)
            assert rel <= 1e-5, (prec, rel)


def test_vjp_single_leaf_equals_direct_sum_adjoint():
    """With all bodies in one leaf the FMM is the direct sum: the VJP must equal
    the closed-form adjoint of the direct sum."""
    d = ic.uniform(60, seed=402, hi=10.0)
    ctx, b = _ctx(d, 4, 0.1, fmm.ALGO_REFERENCE, n_max=64, bounds=(0.0, 10.0, 0.0, 10.0))
    r = ic.SplitMix64(3)
    axb, ayb = r.normal(60), r.normal(60)
    xb, yb, mb = ctx.vjp(b, axb, ayb)
    ox, oy, om = O.n2_vjp(b.x, b.y, b.mass, 0.1, axb, ayb)
    np.testing.assert_allclose(xb, ox, rtol=1e-11, atol=1e-11 * np.abs(ox).max())
    np.testing.assert_allclose(yb, oy, rtol=1e-11, atol=1e-11 * np.abs(oy).max())
    np.testing.assert_allclose(mb, om, rtol=1e-11, atol=1e-11 * np.abs(om).max())


@pytest.mark.parametrize("p,bound", [(6, 1e-3), (8, 1e-4)])
def test_vjp_corrected_close_to_direct_sum_adjoint_8192(p, bound):
    """At N = 8192 the corrected variant's VJP approximates the adjoint of the
    FP64 direct sum (same tolerance class as its forces)."""
    d = ic.plummer(8192, seed=403)
    eps = 0.05
    ctx, b = _ctx(d, p, eps, fmm.ALGO_CORRECTED)
    _forward(ctx, b)
    r = ic.SplitMix64(4)
    axb, ayb = r.normal(len(b)), r.normal(len(b))
    xb, yb, mb = ctx.vjp(b, axb, ayb)
    ox, oy, om = O.n2_vjp(b.x, b.y, b.mass, eps, axb, ayb, threads=8)
    for g, o in ((xb, ox), (yb, oy), (mb, om)):
        rel = np.linalg.norm(g - o) / np.linalg.norm(o)
        assert rel < bound, rel


def test_autograd_function_and_checkpointed_leapfrog():
    torch = pytest.importorskip("torch")
    from galax_amd import autograd as GA

    d = ic.plummer(2048, seed=404)
    eps = 0.05
    ctx, b = _ctx(d, 4, eps, fmm.ALGO_CORRECTED)
    dev = torch.device("cuda:0")
    x = torch.tensor(b.x, device=dev, requires_grad=True)
    y = torch.tensor(b.y, device=dev, requires_grad=True)
    m = torch.tensor(b.mass, device=dev, requires_grad=True)
    ax, ay = GA.fmm_forces(x, y, m, ctx)
    r = ic.SplitMix64(5)
    wx = torch.tensor(r.normal(len(b)), device=dev)
    wy = torch.tensor(r.normal(len(b)), device=dev)
    loss = (ax * wx + ay * wy).sum()
    gx, gy, gm = torch.autograd.grad(loss, (x, y, m))
    xb, yb, mb = ctx.vjp(b, wx.cpu().numpy(), wy.cpu().numpy())
    np.testing.assert_allclose(gx.cpu().numpy(), xb, rtol=1e-12, atol=1e-12 * np.abs(xb).max())
    np.testing.assert_allclose(gm.cpu().numpy(), mb, rtol=1e-12, atol=1e-12 * np.abs(mb).max())

    # 20-step leapfrog, loss = MSE to a perturbed target, checkpoint every 5 steps;
    # gradient w.r.t. initial x of two bodies vs central differences
    vx = torch.tensor(b.vx, device=dev)
    vy = torch.tensor(b.vy, device=dev)
    tx = torch.tensor(b.x + 0.01 * r.normal(len(b)), device=dev)
    ty = torch.tensor(b.y + 0.01 * r.normal(len(b)), device=dev)
    dt, steps = 1e-3, 20

    def run(x0, with_grad):
        with torch.set_grad_enabled(with_grad):
            xf, yf, _, _ = GA.leapfrog(x0, y.detach(), vx, vy, m.detach(), ctx, dt, steps, checkpoint_every=5)
            return ((xf - tx) ** 2 + (yf - ty) ** 2).mean()

    x0 = x.detach().clone().requires_grad_(True)
    L = run(x0, True)
    (g0,) = torch.autograd.grad(L, (x0,))
    h = 1e-6
    for k in (3, 1000):
        xp = x0.detach().clone()
        xp[k] += h
        xm = x0.detach().clone()
        xm[k] -= h
        fd = (run(xp, False) - run(xm, False)).item() / (2 * h)
        assert fd == pytest.approx(g0[k].item(), rel=1e-4, abs=1e-9)
