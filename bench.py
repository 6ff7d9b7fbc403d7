"""Benchmark of the FMM force evaluation on one MI355X.

    python bench.py [--gpus 1] [--steps 10] [--warmup 3] [--workload plummer1m] [--p 6]

A step is one force evaluation (compute_fmm_force, galax-core/src/lib.rs:1466-1499)
of the whole workload on a static tree, with positions, masses and outputs
resident in HBM. Prints one JSON line (value = bodies/s of the whole job).
The tree and list build is timed separately (t_build_s). The CPU baseline is
the oracle (the C++ restatement of galax-core, OpenMP over its Rayon regions)
timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "N=1M FMM forces/s: FP32-MFMA vs FP64 GPU at ≤1e-5 RMS force err; VJP/fwd time"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def _num_moments(p):
    return (p + 1) * (p + 2) // 2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="plummer1m")
    ap.add_argument("--p", type=int, default=6)
    ap.add_argument("--algo", default="reference", choices=["reference", "corrected"])
    ap.add_argument("--theta", type=float, default=0.3)
    ap.add_argument("--err-samples", type=int, default=10_000)
    ap.add_argument("--cpu-sample-n", type=int, default=1 << 18)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)")

    import torch

    from galax_amd import fmm, ic

    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    d, spec = ic.make(args.workload)
    n = len(d["x"])
    algo = fmm.ALGO_CORRECTED if args.algo == "corrected" else fmm.ALGO_REFERENCE
    cfg = fmm.GpuConfig(max_n=n, p=args.p, eps=0.1, n_max=spec["n_max"], bounds=spec["bounds"], algo=algo,
                        theta=args.theta)
    ctx = fmm.GpuContext(cfg)
    b = fmm.BodiesSoA.from_dict(d)
    log(f"[bench] workload={args.workload} N={n} p={args.p} algo={args.algo}: building tree")
    t0 = time.perf_counter()
    ctx.build_tree(b)
    t_build = time.perf_counter() - t0
    info = ctx.tree_info()
    eps = fmm.default_softening(b.x, b.y)  # galax-cli/src/main.rs:182-190, on the sorted bodies
    ctx.set_softening(eps)
    log(f"[bench] tree {info} in {t_build:.3f}s, softening={eps:.6g}")

    x = torch.from_numpy(b.x).to(dev)
    y = torch.from_numpy(b.y).to(dev)
    m = torch.from_numpy(b.mass).to(dev)
    ax = torch.empty_like(x)
    ay = torch.empty_like(x)
    phi = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.evaluate_device(x, y, m, ax, ay, phi)  # first call also builds the static M2L key tables
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    for _ in range(args.warmup):
        ctx.evaluate_device(x, y, m, ax, ay, phi)
    torch.cuda.synchronize()
    passes = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.evaluate_device(x, y, m, ax, ay, phi)
        tm = ctx.timings()
        for k, v in tm.passes_ms.items():
            passes[k] = passes.get(k, 0.0) + v
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t0
    ms_per_step = 1e3 * t_total / args.steps
    passes = {k: v / args.steps for k, v in passes.items()}
    log(f"[bench] {ms_per_step:.3f} ms/step; passes (ms) {passes}")

    # error vs the FP64 direct sum on sampled targets (device direct sum)
    axh, ayh = ax.cpu().numpy(), ay.cpu().numpy()
    rs = ic.SplitMix64(4242)
    samples = np.unique((rs.next_u64(args.err_samples) % np.uint64(n)).astype(np.uint64))
    ex, ey, _ = fmm.compute_n2_force(b.x, b.y, b.mass, eps, targets=samples)
    si = samples.astype(np.int64)
    mag = np.maximum(np.hypot(ex, ey), 1e-300)
    vec = np.hypot(axh[si] - ex, ayh[si] - ey) / mag
    magerr = np.abs(np.hypot(axh[si], ayh[si]) - np.hypot(ex, ey)) / np.maximum(np.hypot(ex, ey), 1e-20)

    # roofline of the dominant kernel (algorithmic flops / bytes per launch)
    S = _num_moments(args.p)
    I = tm.p2p_interactions
    n_leaves = info["n_leaves"]
    dominant = max(("l2p_p2p", "m2l", "p2m", "m2m", "l2l"), key=lambda k: passes.get(k, 0.0))
    flops = {
        "l2p_p2p": 17.0 * I + 4.0 * S * n,  # near field incl. potential + local evaluation
        "m2l": 2.0 * S * S * tm.m2l_pairs_executed,
    }
    bytes_ = {
        "p2m": 24.0 * n + n_leaves * S * 8.0,
        "m2m": 2.0 * info["n_nodes"] * S * 8.0,
        "l2l": 2.0 * n_leaves * S * 8.0,
    }
    t_dom = passes[dominant] * 1e-3
    if dominant in flops:
        achieved = flops[dominant] / t_dom / 1e12
        roof = {"kernel": dominant, "bound": "valu" if dominant == "l2p_p2p" else "mfma", "achieved": achieved,
                "peak": FP64_PEAK_TFS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFS, "traffic": None}
    else:
        achieved = bytes_[dominant] / t_dom / 1e9
        roof = {"kernel": dominant, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, spec, log)

    out = {
        "metric": METRIC,
        "value": n / (ms_per_step * 1e-3),
        "unit": "bodies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 Plummer sphere, 2D)",
        "config": {"workload": args.workload, "N": n, "p": args.p, "n_max": spec["n_max"], "algo": args.algo,
                   "precision": "fp64", "softening": eps, "bounds": list(spec["bounds"]),
                   "theta": args.theta if args.algo == "corrected" else None},
        "roofline": roof,
        "cpu_baseline": cpu,
        "passes_ms": passes,
        "t_build_s": t_build,
        "t_first_eval_s": t_first,
        "tree": info,
        "p2p_interactions": I,
        "p2p_interactions_per_s": I / (passes["l2p_p2p"] * 1e-3),
        "m2l_pairs_executed": tm.m2l_pairs_executed,
        "m2l_pairs_reference": tm.m2l_pairs_reference,
        "m2l_keys": tm.m2l_keys,
        "err_vs_direct": {"samples": int(len(samples)), "rms_vec": float(np.sqrt(np.mean(vec ** 2))),
                          "max_vec": float(vec.max()), "max_mag": float(magerr.max()),
                          "mean_mag": float(magerr.mean())},
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(args, spec, log):
    """galax-core's compute_fmm_force as restated by the oracle, OpenMP threads
    over the reference's Rayon regions, on a bounded sample of the workload
    (same generator, fewer bodies); 1 warm-up, median of 5."""
    from galax_amd import ic
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    nsamp = min(args.cpu_sample_n, spec.get("n", 1 << 30))
    d, sp = ic.make(args.workload, n=nsamp)
    od = {k: np.array(d[k], np.float64) for k in ("x", "y", "vx", "vy", "m", "ax", "ay")}
    t0 = time.perf_counter()
    t = O.tree_build(od, sp["n_max"], sp["bounds"])
    eps = O.default_softening(od["x"], od["y"])
    if args.algo == "corrected":
        m2l, p2p = O.corrected_lists(t, od["x"], od["y"], args.theta)
        algo = 1
    else:
        m2l, p2p = O.m2l_lists(t), O.p2p_lists(t)
        algo = 0
    t_setup = time.perf_counter() - t0
    times = []
    for _ in range(6):
        t0 = time.perf_counter()
        O.compute_fmm_force(od["x"], od["y"], od["m"], t, m2l, p2p, args.p, eps, algo=algo, threads=threads)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times[1:]))
    log(f"[bench] cpu baseline: N={nsamp} setup {t_setup:.2f}s eval median {med:.3f}s on {threads} threads")
    return {"value": nsamp / med, "unit": "bodies/s", "cores": threads, "kind": "port",
            "sample": f"{args.workload} generator at N={nsamp}, p={args.p}, n_max={sp['n_max']}; "
                      f"compute_fmm_force median of 5 after 1 warm-up ({med:.3f} s); "
                      f"tree+lists setup {t_setup:.2f} s not included"}


if __name__ == "__main__":
    main()
