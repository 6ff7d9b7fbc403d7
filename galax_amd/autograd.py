"""torch.autograd wrappers over the C ABI of libgalax_fmm.so.

`fmm_forces(x, y, m, ctx)` returns the FMM accelerations (ax, ay) of FP64
tensors (CPU or CUDA) in the tree's body order; its backward calls the
library's VJP (galax_fmm_vjp / galax_fmm_vjp_device), which recomputes the
forward expansions and runs the adjoint of every pass on the static tree.

`leapfrog(...)` is galax-integrate's kick-drift-kick (galax-integrate/src/lib.rs:158-189)
written with these forces, differentiable, checkpointed every k steps with
torch.utils.checkpoint (activations of a segment are recomputed in backward).
"""
from __future__ import annotations

import numpy as np
import torch
import torch.utils.checkpoint as tcp

from . import fmm


class FmmForce(torch.autograd.Function):
    @staticmethod
    def forward(fctx, x, y, m, gctx: fmm.GpuContext):
        if x.dtype != torch.float64:
            raise TypeError("fmm_forces works in float64")
        xs, ys, ms = x.detach().clone(), y.detach().clone(), m.detach().clone()
        if x.is_cuda:
            ax, ay = torch.empty_like(xs), torch.empty_like(xs)
            # the library runs on its own stream: torch's pending work must be done
            torch.cuda.current_stream(x.device).synchronize()
            gctx.evaluate_device(xs, ys, ms, ax, ay, None)
        else:
            b = fmm.BodiesSoA(xs.numpy(), ys.numpy(), np.zeros(len(xs)), np.zeros(len(xs)), ms.numpy(),
                              np.zeros(len(xs)), np.zeros(len(xs)), None)
            gctx.evaluate(b)
            ax, ay = torch.from_numpy(b.ax), torch.from_numpy(b.ay)
        fctx.save_for_backward(xs, ys, ms)
        fctx.gctx = gctx
        return ax, ay

    @staticmethod
    def backward(fctx, gax, gay):
        xs, ys, ms = fctx.saved_tensors
        gctx = fctx.gctx
        gax = gax.contiguous() if gax is not None else torch.zeros_like(xs)
        gay = gay.contiguous() if gay is not None else torch.zeros_like(xs)
        if xs.is_cuda:
            x2, y2, m2 = xs.clone(), ys.clone(), ms.clone()
            xb, yb, mb = torch.empty_like(xs), torch.empty_like(xs), torch.empty_like(xs)
            torch.cuda.current_stream(xs.device).synchronize()
            gctx.vjp_device(x2, y2, m2, gax, gay, xb, yb, mb)
        else:
            b = fmm.BodiesSoA(xs.numpy().copy(), ys.numpy().copy(), np.zeros(len(xs)), np.zeros(len(xs)),
                              ms.numpy().copy(), np.zeros(len(xs)), np.zeros(len(xs)), None)
            xb, yb, mb = (torch.from_numpy(a) for a in gctx.vjp(b, gax.numpy(), gay.numpy()))
        return xb, yb, mb, None


def fmm_forces(x, y, m, gctx: fmm.GpuContext):
    return FmmForce.apply(x, y, m, gctx)


def _kdk_segment(gctx, m, dt, k, x, y, vx, vy, ax, ay):
    half = 0.5 * dt
    for _ in range(k):
        vx = vx + ax * half
        vy = vy + ay * half
        x = x + vx * dt
        y = y + vy * dt
        ax, ay = fmm_forces(x, y, m, gctx)
        vx = vx + ax * half
        vy = vy + ay * half
    return x, y, vx, vy, ax, ay


def leapfrog(x, y, vx, vy, m, gctx: fmm.GpuContext, dt: float, n_steps: int, checkpoint_every: int = 10):
    """n_steps KDK steps (tree static), differentiable in x, y, vx, vy, m.
    Segments of `checkpoint_every` steps keep only their inputs; their
    intermediate states are recomputed during backward. checkpoint_every = 0
    turns checkpointing off (one segment, every activation kept)."""
    if checkpoint_every < 0:
        raise ValueError(f"checkpoint_every must be >= 0, got {checkpoint_every}")
    if n_steps < 0:
        raise ValueError(f"n_steps must be >= 0, got {n_steps}")
    ax, ay = fmm_forces(x, y, m, gctx)
    done = 0
    while done < n_steps:
        k = min(checkpoint_every, n_steps - done) if checkpoint_every > 0 else n_steps - done
        if checkpoint_every > 0 and torch.is_grad_enabled():
            x, y, vx, vy, ax, ay = tcp.checkpoint(_kdk_segment, gctx, m, dt, k, x, y, vx, vy, ax, ay,
                                                  use_reentrant=False)
        else:
            x, y, vx, vy, ax, ay = _kdk_segment(gctx, m, dt, k, x, y, vx, vy, ax, ay)
        done += k
    return x, y, vx, vy
