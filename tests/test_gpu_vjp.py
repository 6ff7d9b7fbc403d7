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
