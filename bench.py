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
FP32_PEAK_TFS = 157.3


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
    ap.add_argument("--cpu-sample-n", type=int, default=0,
                    help="bodies of the CPU baseline (0: the whole workload)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--vjp-steps", type=int, default=3)
    ap.add_argument("--forms", default="fp64,fp64_exact,mixed_fp32,mixed_bf16x3,corrected_fp64,corrected_mixed_fp32",
                    help="comma-separated forms to run (the headline form and fp64 are always run)")
    ap.add_argument("--corrected", default="8:0.35",
                    help="p:theta of the corrected_* forms (the corrected variant at an order and opening angle "
                         "that bring the forces within 1e-5 RMS of the direct sum on plummer1m, with a margin of "
                         "about 2 on three independent samples)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="after the forms, copy a buffer of this many MiB on the device (a kernel of known "
                         "traffic that scales hardware byte counters)")
    ap.add_argument("--form", default="mixed_fp32", choices=["fp64", "fp64_exact", "mixed_fp32", "mixed_bf16x3"],
                    help="the form whose throughput is `value`")
    ap.add_argument("--config", default=None, choices=["c1", "c2", "c3", "c4", "c5"],
                    help="instead of the headline line, measure one of the workload configurations "
                         "(uniform10k, plummer1m p=4 + 100 KDK steps, disk4m p-sweep, galaxy262k VJP, "
                         "clustered16m) and print its JSON line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        sys.exit("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)")
    if args.config:
        import bench_configs
        print(json.dumps(bench_configs.run(args.config, args)), flush=True)
        return

    import torch

    from galax_amd import fmm, ic

    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    d, spec = ic.make(args.workload)
    n = len(d["x"])
    algo = fmm.ALGO_CORRECTED if args.algo == "corrected" else fmm.ALGO_REFERENCE
    cp, ctheta = args.corrected.split(":")
    cp, ctheta = int(cp), float(ctheta)
    # (name, precision, algo, p, theta, the form its speedup and difference are taken against)
    all_forms = [("fp64", fmm.PREC_FP64, algo, args.p, args.theta, None),
                 ("fp64_exact", fmm.PREC_FP64_EXACT, algo, args.p, args.theta, "fp64"),
                 ("mixed_fp32", fmm.PREC_MIXED_FP32, algo, args.p, args.theta, "fp64"),
                 ("mixed_bf16x3", fmm.PREC_MIXED_BF16X3, algo, args.p, args.theta, "fp64"),
                 ("corrected_fp64", fmm.PREC_FP64, fmm.ALGO_CORRECTED, cp, ctheta, None),
                 ("corrected_mixed_fp32", fmm.PREC_MIXED_FP32, fmm.ALGO_CORRECTED, cp, ctheta, "corrected_fp64")]
    headline = args.form
    wanted = set(args.forms.split(",")) | {"fp64", headline}
    if "corrected_mixed_fp32" in wanted:
        wanted.add("corrected_fp64")
    forms = [f for f in all_forms if f[0] in wanted]
    base_of = {f[0]: f[5] for f in forms}
    res = {}
    eps = None
    for name, prec, falgo, fp, ftheta, _ in forms:
        cfg = fmm.GpuConfig(max_n=n, p=fp, eps=0.1, n_max=spec["n_max"], bounds=spec["bounds"], algo=falgo,
                            theta=ftheta, precision=prec)
        ctx = fmm.GpuContext(cfg)
        b = fmm.BodiesSoA.from_dict(d)
        log(f"[bench] {name}: workload={args.workload} N={n} p={fp} algo={falgo} theta={ftheta}: building tree")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.build_tree(b)
        t_build = time.perf_counter() - t0
        build_ms = ctx.build_timings()
        info = ctx.tree_info()
        if eps is None:
            eps = fmm.default_softening(b.x, b.y)  # galax-cli/src/main.rs:182-190, on the sorted bodies
        ctx.set_softening(eps)
        x = torch.from_numpy(b.x).to(dev)
        y = torch.from_numpy(b.y).to(dev)
        m = torch.from_numpy(b.mass).to(dev)
        ax, ay, phi = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.evaluate_device(x, y, m, ax, ay)  # the first call also builds the static M2L key tables
        torch.cuda.synchronize()
        t_first = time.perf_counter() - t0
        # the tensors were checked by that call: the timed loops pass their
        # device pointers (raw pointers skip the per-call checks)
        px, py, pm, pax, pay, pphi = (t.data_ptr() for t in (x, y, m, ax, ay, phi))
        for _ in range(args.warmup):
            ctx.evaluate_device(px, py, pm, pax, pay)
        torch.cuda.synchronize()
        # the timed step is compute_fmm_force's output: accelerations only,
        # without the per-pass timing events
        ctx.set_timing(False)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.evaluate_device(px, py, pm, pax, pay)
        torch.cuda.synchronize()
        t_total = time.perf_counter() - t0
        ms = 1e3 * t_total / args.steps
        # the per-pass breakdown, in separate steps with the events on
        ctx.set_timing(True)
        passes = {}
        for _ in range(args.steps):
            ctx.evaluate_device(px, py, pm, pax, pay)
            tm = ctx.timings()
            for k, v in tm.passes_ms.items():
                passes[k] = passes.get(k, 0.0) + v
        passes = {k: v / args.steps for k, v in passes.items()}
        log(f"[bench] {name}: {ms:.3f} ms/step (build {t_build:.3f}s, first eval {t_first:.3f}s); passes {passes}")
        # the same step with the potential as well (an extra output of this library)
        nphi = max(1, args.steps // 2)
        ctx.evaluate_device(x, y, m, ax, ay, phi)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(nphi):
            ctx.evaluate_device(px, py, pm, pax, pay, pphi)
        torch.cuda.synchronize()
        ms_phi = 1e3 * (time.perf_counter() - t0) / nphi
        log(f"[bench] {name}: with the potential {ms_phi:.3f} ms/step")
        vjp_ms = None
        if name in ("fp64", headline, "corrected_fp64", "corrected_mixed_fp32") and args.vjp_steps > 0:
            # VJP of the accelerations (forward recomputed inside, adjoint passes in FP64)
            g = torch.Generator(device="cpu").manual_seed(7)
            axb = torch.randn(n, generator=g, dtype=torch.float64).to(dev)
            ayb = torch.randn(n, generator=g, dtype=torch.float64).to(dev)
            xb, yb, mb = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
            ctx.vjp_device(x, y, m, axb, ayb, xb, yb, mb)
            torch.cuda.synchronize()
            vp = [t.data_ptr() for t in (x, y, m, axb, ayb, xb, yb, mb)]
            t0 = time.perf_counter()
            for _ in range(args.vjp_steps):
                ctx.vjp_device(*vp)
            torch.cuda.synchronize()
            vjp_ms = 1e3 * (time.perf_counter() - t0) / args.vjp_steps
            log(f"[bench] {name}: vjp {vjp_ms:.3f} ms ({vjp_ms / ms:.2f}x the forward)")
            del axb, ayb, xb, yb, mb
        res[name] = dict(ms=ms, ms_phi=ms_phi, passes=passes, t_build=t_build, t_first=t_first, tm=tm, info=info,
                         ax=ax.cpu().numpy(), ay=ay.cpu().numpy(), bodies=b, vjp_ms=vjp_ms, build_ms=build_ms)
        del ctx, x, y, m, ax, ay, phi
        torch.cuda.empty_cache()

    if args.copy_probe_mb > 0:
        src = torch.empty(args.copy_probe_mb << 20, dtype=torch.uint8, device=dev).fill_(1)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {args.copy_probe_mb} MiB read + {args.copy_probe_mb} MiB written")
        del src, dst

    b = res["fp64"]["bodies"]
    rs = ic.SplitMix64(4242)
    samples = np.unique((rs.next_u64(args.err_samples) % np.uint64(n)).astype(np.uint64))
    si = samples.astype(np.int64)
    ex, ey, _ = fmm.compute_n2_force(b.x, b.y, b.mass, eps, targets=samples)
    emag = np.hypot(ex, ey)
    errs = {}
    for name, r in res.items():
        vec = np.hypot(r["ax"][si] - ex, r["ay"][si] - ey) / np.maximum(emag, 1e-300)
        magerr = np.abs(np.hypot(r["ax"][si], r["ay"][si]) - emag) / np.maximum(emag, 1e-20)
        e = {"rms_vec_vs_direct": float(np.sqrt(np.mean(vec ** 2))), "max_vec_vs_direct": float(vec.max()),
             "validate_max_mag": float(magerr.max()), "validate_mean_mag": float(magerr.mean())}
        base = base_of.get(name)
        if base is not None:  # against the FP64 form of the same variant, p and theta
            ref = res[base]
            d64 = np.hypot(r["ax"] - ref["ax"], r["ay"] - ref["ay"]) / np.maximum(np.hypot(ref["ax"], ref["ay"]),
                                                                                  1e-300)
            e["rms_vs_fp64"] = float(np.sqrt(np.mean(d64 ** 2)))
            e["max_vs_fp64"] = float(d64.max())
            e["speedup_vs_fp64"] = ref["ms"] / r["ms"]
        errs[name] = e
    # the corrected forms' RMS relative error against the direct sum depends on
    # the sample (a few bodies with nearly cancelling forces dominate it): three
    # independent 10,000-draw samples
    per_sample = {}
    for name in ("corrected_fp64", "corrected_mixed_fp32"):
        if name not in res:
            continue
        r, vals = res[name], []
        for seed in (4242, 77, 2026):
            sm = np.unique((ic.SplitMix64(seed).next_u64(10_000) % np.uint64(n)).astype(np.uint64))
            sx, sy, _ = fmm.compute_n2_force(b.x, b.y, b.mass, eps, targets=sm)
            k = sm.astype(np.int64)
            vec = np.hypot(r["ax"][k] - sx, r["ay"][k] - sy) / np.maximum(np.hypot(sx, sy), 1e-300)
            vals.append(float(np.sqrt(np.mean(vec ** 2))))
        per_sample[name] = vals
    log(f"[bench] errors {errs}; corrected per sample {per_sample}")

    # roofline of the headline form's dominant pass (algorithmic flops or bytes per launch)
    H = res[headline]
    stages = stage_rooflines(H, n, args.p, headline, args.algo == "corrected")
    dominant = max(("l2p_p2p", "m2l", "p2m", "m2m", "l2l"), key=lambda k: stages.get(k, {}).get("ms", 0.0))
    st = stages[dominant]
    roof = {"kernel": dominant, "bound": st["bound"], "achieved": st["achieved"], "peak": st["peak"],
            "unit": st["unit"], "frac": st["frac"], "traffic": measured_traffic(args, headline, dominant, spec["n_max"])}
    tm, info = H["tm"], H["info"]
    I = tm.p2p_interactions

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, spec, log)

    out = {
        "metric": METRIC,
        "value": n / (H["ms"] * 1e-3),
        "unit": "bodies/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": H["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if headline.startswith("mixed") else "f64",
        "data": "synthetic (seeded SplitMix64 2D Plummer sphere)",
        "config": {"workload": args.workload, "N": n, "p": args.p, "n_max": spec["n_max"], "algo": args.algo,
                   "form": headline, "softening": eps, "bounds": list(spec["bounds"]),
                   "theta": args.theta if args.algo == "corrected" else None},
        "roofline": roof,
        "cpu_baseline": cpu,
        "forms": {k: {"ms_per_step": r["ms"], "bodies_per_s": n / (r["ms"] * 1e-3),
                      "ms_per_step_with_potential": r["ms_phi"], "passes_ms": r["passes"],
                      "t_build_s": r["t_build"], "t_first_eval_s": r["t_first"],
                      "p2p_interactions_per_s": r["tm"].p2p_interactions / (r["passes"]["l2p_p2p"] * 1e-3),
                      "vjp_ms": r["vjp_ms"], "vjp_over_fwd": r["vjp_ms"] / r["ms"] if r["vjp_ms"] else None,
                      **errs[k]}
                  for k, r in res.items()},
        "stages": stages,
        "build_ms": H["build_ms"],
        "tree": info,
        "m2l_tiles": {"groups": tm.m2l_groups, "steps": tm.m2l_steps, "items": tm.m2l_items},
        "p2p_interactions": I,
        "m2l_pairs_executed": tm.m2l_pairs_executed,
        "m2l_pairs_reference": tm.m2l_pairs_reference,
        "m2l_keys": tm.m2l_keys,
        "vjp_over_fwd": H["vjp_ms"] / H["ms"] if H["vjp_ms"] else None,
    }
    if "corrected_mixed_fp32" in res:
        # the corrected variant meets the error target against the direct sum
        # itself (the reference variant's own error is set by its algorithm)
        c64, c32 = res["corrected_fp64"], res["corrected_mixed_fp32"]
        out["corrected_at_1e-5_vs_direct"] = {
            "p": cp, "theta": ctheta, "fp64_ms": c64["ms"], "mixed_fp32_ms": c32["ms"],
            "speedup_mixed_vs_fp64": c64["ms"] / c32["ms"],
            "rms_vec_vs_direct": {"fp64": errs["corrected_fp64"]["rms_vec_vs_direct"],
                                  "mixed_fp32": errs["corrected_mixed_fp32"]["rms_vec_vs_direct"]},
            "rms_vec_vs_direct_per_sample": {"samples": "10,000 SplitMix64 draws each, seeds 4242, 77, 2026",
                                             "fp64": per_sample.get("corrected_fp64"),
                                             "mixed_fp32": per_sample.get("corrected_mixed_fp32")},
            "vjp_over_fwd": {"fp64": c64["vjp_ms"] / c64["ms"] if c64["vjp_ms"] else None,
                             "mixed_fp32": c32["vjp_ms"] / c32["ms"] if c32["vjp_ms"] else None}}
    print(json.dumps(out), flush=True)


def source_sha() -> str:
    """SHA-256 of the HIP sources (galax_amd/csrc/*.hip, *.h, sorted by name):
    the build a traffic measurement belongs to."""
    import hashlib

    h = hashlib.sha256()
    cdir = os.path.join(ROOT, "galax_amd", "csrc")
    for name in sorted(os.listdir(cdir)):
        if name.endswith((".hip", ".h")):
            h.update(name.encode())
            with open(os.path.join(cdir, name), "rb") as f:
                h.update(f.read())
    return h.hexdigest()


def measured_traffic(args, form, stage, n_max):
    """HBM bytes per launch of the stage's main kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (profiles/traffic.json), when they
    were taken on this workload, p and form with these very sources (the
    file's source_sha); else None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if (t.get("workload"), t.get("p"), t.get("form")) != (args.workload, args.p, form):
        return None
    if t.get("source_sha") != source_sha():
        return None  # measured on other kernel sources: stale
    mode = {"fp64": 0, "mixed_fp32": 1, "mixed_bf16x3": 2}.get(form)
    prec = "float" if form.startswith("mixed") else "double"
    sym = form.startswith("mixed") and n_max <= 32 and os.environ.get("GALAX_NEAR_GATHER") != "1"
    # the timed step's kernels: accelerations only (no potential)
    name = {"m2l": f"gx::k_m2l_tiles<{args.p}, {mode}, false>",
            "l2p_p2p": f"gx::k_p2p_sym<{args.p}, false>" if sym else f"gx::k_l2p_p2p<{args.p}, {prec}, false>"}.get(stage)
    k = t["kernels"].get(name)
    return k["traffic_bytes"] if k else None


def stage_rooflines(r, n, p, form, algo_corrected=False):
    """Achieved rate of each pass and build stage against the MI355X peak that
    bounds it, from algorithmic bytes or flops per launch (N bodies, L leaves,
    V nodes, S coefficients, E executed M2L pairs, I P2P interactions; w = 8:
    expansions and bodies are FP64 in every form)."""
    passes, tm, info, b = r["passes"], r["tm"], r["info"], r["build_ms"]
    S = _num_moments(p)
    L, V, E, I = info["n_leaves"], info["n_nodes"], tm.m2l_pairs_executed, tm.p2p_interactions
    mixed = form.startswith("mixed")
    out = {}

    def hbm(name, ms, nbytes):
        if ms and ms > 0:
            a = nbytes / (ms * 1e-3) / 1e9
            out[name] = {"ms": ms, "bound": "hbm", "bytes": nbytes, "achieved": a, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": a / HBM_PEAK_GBS}

    def comp(name, ms, flops, bound, peak):
        if ms and ms > 0:
            a = flops / (ms * 1e-3) / 1e12
            out[name] = {"ms": ms, "bound": bound, "flops": flops, "achieved": a, "peak": peak, "unit": "TFLOP/s",
                         "frac": a / peak}

    hbm("sort", b["sort"], 24.0 * n + 6 * 2 * 12.0 * n)      # Morton keys + 6 radix passes of (key, index)
    hbm("permute", b["permute"], 2 * 56.0 * n)               # gather of the 7 body arrays
    hbm("tree", b["tree"], 24.0 * n + V * 2 * 64.0)          # leaf geometry + parent rounds
    hbm("p2m", passes.get("p2m"), 24.0 * n + L * S * 8.0)
    hbm("m2m", passes.get("m2m"), 2.0 * L * S * 8.0)         # reference variant: level-1 parents of the leaves
    if algo_corrected:  # the reference variant's one L2L level runs inside the L2P+P2P kernel
        hbm("l2l", passes.get("l2l"), 2.0 * V * S * 8.0)
    # the main kernel of each pass alone when it was timed (no combine pass)
    comp("m2l", passes.get("m2l_kernel") or passes.get("m2l"), 2.0 * S * S * E, "mfma",
         FP32_PEAK_TFS if mixed else FP64_PEAK_TFS)
    comp("l2p_p2p", passes.get("p2p_kernel") or passes.get("l2p_p2p"), 14.0 * I + 4.0 * S * n, "valu",
         FP32_PEAK_TFS if mixed else FP64_PEAK_TFS)
    return out


def cpu_baseline(args, spec, log):
    """galax-core's compute_fmm_force as restated by the oracle, OpenMP threads
    over the reference's Rayon regions, on the full workload; 1 warm-up,
    median of 3. The tree and lists are the device build's, exported and
    imported into the oracle: they are bit-identical to the oracle's own
    (the parity tests check that), whose reference-rule list build alone takes
    about a minute on one core at 2^20 bodies. With --cpu-sample-n below the
    workload size, a smaller sample of the same generator with the oracle's
    own tree and lists instead."""
    from galax_amd import fmm, ic
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    full_n = spec.get("n", 1 << 30)
    nsamp = min(args.cpu_sample_n, full_n) if args.cpu_sample_n > 0 else full_n
    d, sp = ic.make(args.workload, n=nsamp)
    od = {k: np.array(d[k], np.float64) for k in ("x", "y", "vx", "vy", "m", "ax", "ay")}
    t0 = time.perf_counter()
    algo = 1 if args.algo == "corrected" else 0
    if nsamp == full_n and algo == 0:
        ctx = fmm.GpuContext(fmm.GpuConfig(max_n=nsamp, p=args.p, eps=0.1, n_max=sp["n_max"], bounds=sp["bounds"]))
        b = fmm.BodiesSoA.from_dict(d)
        ctx.build_tree(b)
        tree, m2l_csr, p2p_csr = ctx.export_tree()
        del ctx
        od = {"x": b.x.copy(), "y": b.y.copy(), "m": b.mass.copy()}
        t = O.Tree.from_arrays(tree, sp["n_max"], sp["bounds"])
        m2l, p2p = O.Lists.from_csr(*m2l_csr), O.Lists.from_csr(*p2p_csr)
        how = "tree and lists from the device build (bit-identical to the oracle's)"
    else:
        t = O.tree_build(od, sp["n_max"], sp["bounds"])
        if algo:
            m2l, p2p = O.corrected_lists(t, od["x"], od["y"], args.theta)
        else:
            m2l, p2p = O.m2l_lists(t), O.p2p_lists(t)
        how = "the oracle's own tree and lists"
    eps = O.default_softening(od["x"], od["y"])
    t_setup = time.perf_counter() - t0
    times = []
    for _ in range(4):
        t0 = time.perf_counter()
        O.compute_fmm_force(od["x"], od["y"], od["m"], t, m2l, p2p, args.p, eps, algo=algo, threads=threads)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times[1:]))
    log(f"[bench] cpu baseline: N={nsamp} setup {t_setup:.2f}s eval median {med:.3f}s on {threads} threads")
    return {"value": nsamp / med, "unit": "bodies/s", "cores": threads, "kind": "port",
            "sample": f"{args.workload} at N={nsamp}, p={args.p}, n_max={sp['n_max']}; compute_fmm_force "
                      f"median of 3 after 1 warm-up ({med:.3f} s) on {how}; setup {t_setup:.2f} s not included"}


if __name__ == "__main__":
    main()
