#!/usr/bin/env python3
"""Monte Carlo solves/sec of the Hexa8 cantilever on one MI355X.

Workload (one "step"): the batched mixed-precision Monte Carlo job of
1024 samples whose cube-mesh sizes are drawn uniformly from
n in {16, 24, 32, 40, 48} elements per side; per sample a lognormal
Karhunen-Loeve modulus field (40 terms, exponential covariance, correlation
length 0.25, mean 2.1e5, COV 0.3) evaluated at element centroids, FP32
element integration + coloured assembly, FP32 Jacobi-PCG inner solves under
FP64 iterative refinement (FP64 residual f - K u).  Inputs (meshes, KL
tables, KL coefficients, loads) are resident on the device before the timed
region; each step re-evaluates the modulus, re-assembles and re-solves every
sample.  The PCG preconditioner is a geometric multigrid V-cycle per sample
(--precond mg, default; coarse levels re-assembled every step) or Jacobi
(--precond jacobi, the reference's SolverPCG choice); the other one is timed
beside it for comparison.

Reported beside the headline: the same job in plain FP64 (batched), the
single-sample FP64 GPU PCG, each sample's error against a tightly converged
FP64 solve, the SpMV kernel's roofline (HIP events) and a CPU baseline (the
oracle's CSR/Jacobi restatement of the reference loop, one sample per mesh
size, one core).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MC solves/sec at 1e-8 rel. err, batched mixed vs plain FP64, 1 MI355X"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--samples", type=int, default=1024)
    p.add_argument("--sizes", type=str, default="16,24,32,40,48")
    p.add_argument("--cov", type=float, default=0.3)
    p.add_argument("--terms", type=int, default=40)
    p.add_argument("--corr", type=float, default=0.25)
    p.add_argument("--tol", type=float, default=1e-9, help="FP64 residual tolerance |f-Ku|/|f|")
    p.add_argument("--inner-tol", type=float, default=1e-4)
    p.add_argument("--oracle-tol", type=float, default=1e-13)
    p.add_argument("--max-it", type=int, default=20000)
    p.add_argument("--max-outer", type=int, default=30)
    p.add_argument("--graph-iters", type=int, default=16)
    p.add_argument("--tile", type=int, default=1)
    p.add_argument("--period", type=int, default=None,
                   help="replacement check period (iterations; default 2 with multigrid, 8 with Jacobi)")
    p.add_argument("--drop", type=float, default=None,
                   help="FP32 residual drop that triggers a replacement (default 1e-3 with "
                        "multigrid, 0.1 with Jacobi)")
    p.add_argument("--refine", choices=["reliable", "restart"], default="reliable")
    p.add_argument("--precond", choices=["mg", "jacobi"], default="mg",
                   help="PCG preconditioner: multigrid V-cycle or Jacobi")
    p.add_argument("--no-compare", action="store_true",
                   help="skip the timing of the other preconditioner")
    p.add_argument("--only-mixed", action="store_true", help="skip every comparison run")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                   help="only run a device copy of this many MB (HBM counter calibration)")
    p.add_argument("--seed", type=int, default=20261202)
    p.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison runs")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-threads", type=int, default=-1,
                   help="threads of the multi-core CPU baseline (-1: host cores up to 16, "
                        "0: skip it)")
    p.add_argument("--no-single", action="store_true", help="skip the single-sample FP64 runs")
    p.add_argument("--newmark", type=int, default=0, metavar="STEPS",
                   help="instead of the Monte Carlo job, time STEPS Newmark steps of the "
                        "batched dynamics workload (48x48x24 block, 64 samples)")
    p.add_argument("--newmark-samples", type=int, default=64)
    p.add_argument("--newmark-tol", type=float, default=1e-8)
    p.add_argument("--newmark-cold", action="store_true",
                   help="also time cold (zero) starts of every step's PCG")
    return p.parse_args()


def workload(args):
    sizes = [int(s) for s in args.sizes.split(",")]
    rng = np.random.Generator(np.random.Philox(args.seed))
    draw = rng.integers(0, len(sizes), args.samples)
    counts = [int((draw == k).sum()) for k in range(len(sizes))]
    keep = [k for k in range(len(sizes)) if counts[k] > 0]
    return [sizes[k] for k in keep], [counts[k] for k in keep]


def refinement(args, precond):
    """Replacement check period and trigger: a multigrid iteration cuts the
    residual by about 5x, so checks come every 2 iterations and a
    replacement waits for a 1e-3 drop; Jacobi-PCG keeps 8 and 0.1."""
    mg = precond == "mg"
    return (args.period if args.period is not None else (2 if mg else 8),
            args.drop if args.drop is not None else (1e-3 if mg else 0.1))


def make_batch(msa, ctx, meshes, models, counts, precision, tile, table, xi, args, precond=None):
    b = msa.Batch(ctx, meshes, counts, precision, tile)
    precond = precond or args.precond
    if precond == "mg":
        b.set_preconditioner(msa.MSA_PRECOND_MULTIGRID)
    if precision == msa.MSA_MIXED:
        period, drop = refinement(args, precond)
        b.set_refinement(msa.MSA_REFINE_RELIABLE if args.refine == "reliable"
                         else msa.MSA_REFINE_RESTART, period, drop)
    off = 0
    for c, m in enumerate(models):
        b.set_kl(c, table, m.domain, 2.1e5, args.cov)
        b.set_xi(c, xi[off:off + counts[c]])
        b.set_rhs(c, m.f)
        off += counts[c]
    ctx.synchronize()
    return b


def smooth_bytes(b, models, rs):
    """Algorithmic bytes of all finest-level smoothing launches of one solve.
    Per cycle and active sample the four steps read the values four times
    and move 16 vectors (PRE: z gathered, b, D^-1 read, A written; RESID: A,
    b read, B written; POST1: A, b, D^-1 read, B written; POST2: B, b, D^-1,
    A read, z written), plus the shared block-column indices and slice
    pointers once per launch.  Cycles per sample = iterations + replacements
    (one at the start, one after every replacement that does not finish)."""
    total = 0.0
    for c, m in enumerate(models):
        if b.mg_info(c)["levels"] == 0:
            continue
        st = b.stats(c)
        mesh = b.meshes[c]
        nnz, n = mesh.info["nnz"], 3 * mesh.info["blocks"]
        cyc = (st["iters"] + np.maximum(st["outer"], 1)).astype(np.float64)
        total += cyc.sum() * 4 * (rs * nnz + 4 * rs * n)
        idx = 4.0 * mesh.info["block_nnz"] + 4.0 * (mesh.info["blocks"] + 1)
        total += float(cyc.max()) * 4 * idx
    return total


def mf_work(b, models):
    """FP32 operations and algorithmic HBM bytes of all finest-level
    matrix-free smoothing launches of one solve (None unless every class
    with a hierarchy uses them).  Per cycle and active sample four products;
    each does, for every (free node, incident cell) pair, 72 FMA (three
    rows of the unit element matrix times the cell's 24 values) and 3 FMA
    (modulus scaling): 150 flop.  Bytes per product and sample: the moduli
    (8 per element) and 4 vectors of 4*3*nb on average over the four steps
    (as for the stored-value products, without the values); the grid tables
    once per launch."""
    flops = byts = 0.0
    for c, m in enumerate(models):
        info = b.mg_info(c)
        if info["levels"] == 0:
            continue
        if not info["matrix_free"]:
            return None
        st = b.stats(c)
        cyc = (st["iters"] + np.maximum(st["outer"], 1)).astype(np.float64)
        conn = np.asarray(m.conn).reshape(-1, 8)
        free_node = (np.asarray(m.dof).reshape(-1, 3) >= 0).any(axis=1)
        inc = float(free_node[conn].sum())
        nb = float(free_node.sum())
        flops += cyc.sum() * 4 * inc * 150.0
        byts += cyc.sum() * 4 * (8.0 * conn.shape[0] + 4 * 4 * 3 * nb)
        byts += float(cyc.max()) * 4 * 4.0 * (len(free_node) + conn.shape[0] + nb)
    return flops, byts


def kernel_a_bytes(b, models, counts, rs, launches):
    """Algorithmic bytes moved by all SpMV+dot launches of one solve:
    per active sample and iteration the matrix values and the four vectors
    (z, p_old read; p_new, q written); the shared block-column indices and
    slice pointers once per launch per class with active samples."""
    total = 0.0
    for c, m in enumerate(models):
        st = b.stats(c)
        mesh = b.meshes[c]
        nnz, n = mesh.info["nnz"], 3 * mesh.info["blocks"]
        its = st["iters"].astype(np.float64)
        total += its.sum() * (rs * nnz + 4 * rs * n)
        idx = 4.0 * mesh.info["block_nnz"] + 4.0 * (mesh.info["blocks"] + 1)
        total += float(its.max()) * idx
    return total


def bench_newmark(args):
    """Batched Newmark dynamics: 48x48x24 Hexa8 block (1 x 1 x 0.5, clamped
    at x = 0), consistent mass, density 7.85e-9, Newmark beta 0.25 / gamma
    0.5, dt 1e-4, base acceleration 1.0 sin(2 pi 20 t) in Z, lognormal KL
    modulus (COV 0.2), mixed-precision PCG warm-started every step.  Rate =
    samples x steps / s."""
    import msolve_amd as msa
    from msolve_amd.stochastic import kl_table, standard_normal
    ns, nsteps = args.newmark_samples, args.newmark
    ctx = msa.Context(0)
    model = msa.cantilever(48, 48, 24, lz=0.5)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    table = kl_table(list(model.domain), args.corr, args.terms)
    xi = standard_normal(args.seed, ns, args.terms)
    acc = msa.sinusoidal_history(0.05, 1e-4, 0.05, 1.0, nsteps)

    def make(precision):
        b = msa.Batch(ctx, [mesh], [ns], precision, 1)
        if args.precond == "mg":
            b.set_preconditioner(msa.MSA_PRECOND_MULTIGRID)
        if precision == msa.MSA_MIXED:
            period, drop = refinement(args, args.precond)
            b.set_refinement(msa.MSA_REFINE_RELIABLE, period, drop)
        b.set_kl(0, table, model.domain, 2.1e5, 0.2)
        b.set_xi(0, xi)
        b.eval_kl()
        b.set_dynamics(7.85e-9, 0.0, 0.0, 0.25, 0.5, 1e-4)
        b.set_probe(0, model.n_free - 1)
        b.assemble(0.3)
        ctx.synchronize()
        return b

    b = make(msa.MSA_MIXED)
    b.newmark(acc[:min(2, nsteps)], 2, args.newmark_tol, args.max_it, args.graph_iters)
    ctx.synchronize()
    t = time.perf_counter()
    b.newmark(acc, 2, args.newmark_tol, args.max_it, args.graph_iters)
    ctx.synchronize()
    el = time.perf_counter() - t
    its = [b.dynamic_state(0, s)[3] for s in range(ns)]
    hist = [b.history(0, s) for s in range(ns)]
    st = b.stats(0)
    extra = {}
    if args.newmark_cold:
        t = time.perf_counter()
        b.newmark(acc, 2, args.newmark_tol, args.max_it, args.graph_iters, warm_start=0)
        ctx.synchronize()
        elc = time.perf_counter() - t
        extra["cold_sample_steps_per_s"] = ns * nsteps / elc
        extra["cold_iters_per_step_mean"] = float(np.mean(
            [b.dynamic_state(0, s)[3] for s in range(ns)])) / nsteps
    b.close()
    # accuracy: the same steps in FP64, PCG to 1e-12, for the first 4 samples
    b64 = make(msa.MSA_FP64)
    b64.newmark(acc, 2, 1e-12, args.max_it, args.graph_iters)
    errs = [float(np.linalg.norm(hist[s] - b64.history(0, s)) / np.linalg.norm(b64.history(0, s)))
            for s in range(min(4, ns))]
    b64.close()
    out = {
        "metric": "Newmark sample-steps/sec, batched mixed warm-started PCG, 1 MI355X",
        "value": ns * nsteps / el, "unit": "sample-steps/s", "n_gpus": 1, "steps": nsteps,
        "ms_per_step": 1e3 * el / nsteps, "higher_is_better": True, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"batched Newmark: {ns} lognormal-KL samples (COV 0.2) of a "
                               f"48x48x24 Hexa8 block, rho 7.85e-9, beta 0.25, gamma 0.5, "
                               f"dt 1e-4, 20 Hz base acceleration, "
                               f"{'multigrid' if args.precond == 'mg' else 'Jacobi'}-preconditioned "
                               f"PCG to {args.newmark_tol}",
                   "precond": args.precond,
                   "n_free": model.n_free, "all_ok": bool((st["status"] == msa.MSA_OK).all()),
                   "iters_per_step_mean": float(np.mean(its)) / nsteps,
                   "probe_history_rel_err_vs_fp64": max(errs)},
        "extra": extra,
    }
    mesh.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.newmark and args.gpus == 1:
        return bench_newmark(args)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, device="cuda", dtype=torch.float32)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": 4 * n, "launches": 3}))
        return
    if args.only_mixed:
        args.no_fp64 = args.no_cpu = args.no_single = True
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding of the "
                                   "Monte Carlo samples is described in DESIGN.md",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import msolve_amd as msa
    from msolve_amd.stochastic import kl_table, standard_normal

    sizes, counts = workload(args)
    N = sum(counts)
    log(f"[bench] sizes={sizes} counts={counts} samples={N}")
    ctx = msa.Context(0)
    models = [msa.cantilever(n) for n in sizes]
    meshes = [msa.Mesh(ctx, m.xyz, m.conn, m.dof, m.n_free) for m in models]
    table = kl_table([1.0, 1.0, 1.0], args.corr, args.terms)
    xi = standard_normal(args.seed, N, args.terms)
    run_kw = dict(nu=0.3, tol=args.tol, max_it=args.max_it, inner_tol=args.inner_tol,
                  max_outer=args.max_outer, graph_iters=args.graph_iters)

    # ------------------------------------------------ headline: mixed, batched
    bm = make_batch(msa, ctx, meshes, models, counts, msa.MSA_MIXED, args.tile, table, xi, args)
    log(f"[bench] mixed batch: {bm.device_bytes / 1e9:.1f} GB on device")
    for w in range(args.warmup):
        t = time.perf_counter()
        bm.run(**run_kw)
        ctx.synchronize()
        log(f"[bench] warmup {w}: {time.perf_counter() - t:.2f} s")
    ctx.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        bm.run(**run_kw)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    value = N * args.steps / elapsed
    log(f"[bench] mixed: {args.steps} steps in {elapsed:.2f} s -> {value:.1f} solves/s")
    stats = [bm.stats(c) for c in range(len(models))]
    all_ok = all((s["status"] == msa.MSA_OK).all() for s in stats)

    # roofline of the dominant kernel: one more step with HIP events around
    # every launch on the context stream
    bm.set_kernel_timing(True)
    bm.run(**run_kw)
    kt = bm.kernel_timing()
    bm.set_kernel_timing(False)
    log(f"[bench] kernel timing: {json.dumps(kt)}")
    a_launches = kt["spmv_dot"]["launches"]
    a_bytes = kernel_a_bytes(bm, models, counts, 4, a_launches) / max(a_launches, 1)
    a_s = kt["spmv_dot"]["ms"] / 1e3 / max(a_launches, 1)
    roof_a = dict(kernel=f"k_pcg_a<float,{args.tile}> (SpMV + p update + p.q)",
                  launches=a_launches, avg_launch_ms=1e3 * a_s, avg_bytes_per_launch=a_bytes,
                  achieved=a_bytes / a_s / 1e9 if a_s > 0 else 0.0,
                  ms_total=kt["spmv_dot"]["ms"])
    roof_a["frac"] = roof_a["achieved"] / HBM_PEAK_GBS
    # the dominant kernel: the finest level's smoothing products under multigrid
    mf = mf_work(bm, models) if args.precond == "mg" else None
    bound, unit, peak, flops_per_launch = "hbm", "GB/s", HBM_PEAK_GBS, None
    if args.precond == "mg" and kt["multigrid"]["launches"] > 0:
        launches = kt["multigrid"]["launches"]
        avg_s = kt["multigrid"]["ms"] / 1e3 / launches
        if mf is not None:
            flops_per_launch = mf[0] / launches
            avg_bytes = mf[1] / launches
            bound, unit, peak = "mfma", "TFLOP/s", FP32_PEAK_TFLOPS
            kname = (f"k_mg_smooth<float,1,STEP,2> on the finest level: matrix-free FP32 "
                     f"products on box grids (FP32 VALU, no MFMA; priced against the FP32 peak) "
                     f"of the cycle's four steps PRE, RESID, POST1, POST2")
        else:
            avg_bytes = smooth_bytes(bm, models, 4) / launches
            kname = (f"k_mg_smooth<float,{args.tile},STEP> on the finest level (the cycle's four "
                     f"products: PRE, RESID, POST1, POST2)")
    else:
        launches, avg_bytes, avg_s = a_launches, a_bytes, a_s
        kname = roof_a["kernel"]
    if flops_per_launch is not None:
        achieved = flops_per_launch / avg_s / 1e12 if avg_s > 0 else 0.0
    else:
        achieved = avg_bytes / avg_s / 1e9 if avg_s > 0 else 0.0
    # HBM traffic of k_pcg_a: rocprofv3 FETCH_SIZE/WRITE_SIZE passes (scaled by a
    # device-copy calibration) measured bytes / algorithmic bytes, applied to
    # this run's algorithmic bytes per launch (profiles/k_pcg_a_mixed_montecarlo.md)
    traffic, traffic_ratio = None, None
    cpath = os.path.join(ROOT, "profiles", "k_pcg_a_counters.json" if args.precond == "jacobi"
                         else ("k_mg_smooth_mf_counters.json" if mf is not None
                               else "k_mg_smooth_counters.json"))
    if os.path.exists(cpath):
        cj = json.load(open(cpath))
        traffic_ratio = (cj["fetch_bytes"] + cj["write_bytes"]) / (cj["algorithmic_read"] +
                                                                   cj["algorithmic_write"])
        traffic = avg_bytes * traffic_ratio
    u_mixed = [[bm.solution(c, s) for s in range(counts[c])] for c in range(len(models))]
    iters_mixed = [s["iters"].tolist() for s in stats]
    outer_mixed = [s["outer"].tolist() for s in stats]
    bm.close()

    extra = {"k_pcg_a": roof_a, "kernel_timing_ms": {k: v["ms"] for k, v in kt.items()}}
    # ------------------------------------------------ the other preconditioner
    other = "jacobi" if args.precond == "mg" else "mg"
    if not args.no_compare and not args.only_mixed:
        bo = make_batch(msa, ctx, meshes, models, counts, msa.MSA_MIXED, args.tile, table, xi,
                        args, precond=other)
        bo.run(**run_kw)
        ctx.synchronize()
        t = time.perf_counter()
        bo.run(**run_kw)
        ctx.synchronize()
        to = time.perf_counter() - t
        extra[f"mixed_{other}_solves_per_s"] = N / to
        extra[f"mixed_{other}_iters_mean_per_size"] = {
            str(n): float(np.mean(bo.stats(c)["iters"])) for c, n in enumerate(sizes)}
        log(f"[bench] mixed with {other}: {to:.2f} s -> {N / to:.1f} solves/s")
        bo.close()
    # ------------------------------------------------ FP64: batched + oracle
    max_err = None
    if not args.no_fp64:
        b64 = make_batch(msa, ctx, meshes, models, counts, msa.MSA_FP64, args.tile, table, xi, args)
        ctx.synchronize()
        t = time.perf_counter()
        b64.run(**run_kw)
        ctx.synchronize()
        t64 = time.perf_counter() - t
        extra["fp64_batched_solves_per_s"] = N / t64
        extra["fp64_iters_mean_per_size"] = {
            str(n): float(np.mean(b64.stats(c)["iters"])) for c, n in enumerate(sizes)}
        log(f"[bench] fp64 batched: {t64:.2f} s -> {N / t64:.1f} solves/s, iters "
            f"{extra['fp64_iters_mean_per_size']}")
        kw = dict(run_kw)
        kw["tol"] = args.oracle_tol
        b64.run(**kw)
        ctx.synchronize()
        errs = []
        for c in range(len(models)):
            st = b64.stats(c)
            assert (st["status"] == msa.MSA_OK).all(), "FP64 oracle solve did not converge"
            for s in range(counts[c]):
                ur = b64.solution(c, s)
                errs.append(np.linalg.norm(u_mixed[c][s] - ur) / np.linalg.norm(ur))
        max_err = float(np.max(errs))
        extra["rel_err_vs_fp64"] = dict(max=max_err, median=float(np.median(errs)),
                                        oracle_tol=args.oracle_tol)
        log(f"[bench] mixed vs FP64 oracle: max rel err {max_err:.3e}")
        b64.close()

    # ------------------------------------------------ single-sample FP64 GPU PCG
    if not args.no_single:
        per = []
        off = 0
        for c, m in enumerate(models):
            b1 = make_batch(msa, ctx, [meshes[c]], [m], [1], msa.MSA_FP64, 1, table,
                            xi[off:off + 1], args, precond="jacobi")
            b1.run(**run_kw)      # warm-up (graph capture)
            ctx.synchronize()
            t = time.perf_counter()
            b1.run(**run_kw)
            ctx.synchronize()
            per.append(time.perf_counter() - t)
            b1.close()
            off += counts[c]
        t_single = sum(counts[c] * per[c] for c in range(len(models)))
        extra["fp64_single_solves_per_s"] = N / t_single
        extra["fp64_single_s_per_size"] = dict(zip(map(str, sizes), per))
        log(f"[bench] fp64 single-sample: {N / t_single:.1f} solves/s ({per})")

    # ------------------------------------------------ CPU baseline
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        otab = O.kl_table([1.0, 1.0, 1.0], args.corr, args.terms)
        per, t_meas = [], 0.0
        off = 0
        for c, n in enumerate(sizes):
            out = O.mc_reference(n, otab, xi[off:off + 1], cov=args.cov, tol=args.tol,
                                 keep_u=False, storage="csr")
            per.append(out["times"]["all"])
            t_meas += out["times"]["all"]
            off += counts[c]
            log(f"[bench] cpu {n}^3: {out['times']['all']:.2f} s, {int(out['iters'][0])} its")
        t_cpu = sum(counts[c] * per[c] for c in range(len(sizes)))
        cpu = dict(value=N / t_cpu, unit="solves/s", cores=1, kind="port",
                   sample=f"one sample per mesh size {sizes} (first of each class), oracle "
                          f"CSR/Jacobi-PCG FP64 restatement of the reference per-sample loop "
                          f"(KL modulus, Ke, assembly, PCG to {args.tol}), serial; rate = "
                          f"{N} / sum(class count x measured time), {t_meas:.1f} s measured")
        # the same restatement on several host cores at once (samples are
        # independent; the oracle's C calls release the interpreter lock):
        # every thread solves one other sample of each size, the per-size
        # times under that load weight the class counts as above
        nthr = args.cpu_threads
        if nthr < 0:
            nthr = min(16, len(os.sched_getaffinity(0)))
        if nthr > 1:
            from concurrent.futures import ThreadPoolExecutor

            def cpu_job(tid):
                res, o = [], 0
                for c, n in enumerate(sizes):
                    row = o + (1 + tid) % counts[c]
                    r = O.mc_reference(n, otab, xi[row:row + 1], cov=args.cov, tol=args.tol,
                                       keep_u=False, storage="csr")
                    res.append(r["times"]["all"])
                    o += counts[c]
                return res

            t0 = time.perf_counter()
            with ThreadPoolExecutor(nthr) as ex:
                res = list(ex.map(cpu_job, range(nthr)))
            wall = time.perf_counter() - t0
            per_par = [sum(r[c] for r in res) / nthr for c in range(len(sizes))]
            rate = nthr * N / sum(counts[c] * per_par[c] for c in range(len(sizes)))
            extra["cpu_baseline_parallel"] = dict(
                value=rate, unit="solves/s", cores=nthr, kind="port",
                sample=f"{nthr} threads, each one further sample per mesh size {sizes}, the same "
                       f"oracle restatement; rate = {nthr} x {N} / sum(class count x mean "
                       f"per-size time under load), {wall:.1f} s wall")
            log(f"[bench] cpu x{nthr}: {rate:.2f} solves/s ({wall:.1f} s wall)")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": (f"batched mixed-precision MC: {N} samples, unit-cube Hexa8 cantilevers "
                         f"n in {sizes} (counts {counts}), lognormal KL modulus (mean 2.1e5, "
                         f"COV {args.cov}, {args.terms} terms, corr. length {args.corr}), FP32 "
                         + ("multigrid-preconditioned " if args.precond == "mg" else "")
                         + (f"{'PCG' if args.precond == 'mg' else 'Jacobi-PCG'} with FP64 "
                            f"residual replacement (check every "
                            f"{refinement(args, args.precond)[0]} iterations, "
                            f"{refinement(args, args.precond)[1]}x drop)"
                            if args.refine == "reliable" else
                            f"inner Jacobi-PCG (tol {args.inner_tol}) + restarted FP64 refinement")
                         + f" to |f-Ku|/|f| < {args.tol}"),
            "samples": N, "sizes": sizes, "counts": counts, "tile": args.tile,
            "precond": args.precond,
            "graph_iters": args.graph_iters, "all_converged": bool(all_ok),
            "max_rel_err_vs_fp64": max_err,
        },
        "roofline": {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
                     "frac": achieved / peak, "traffic": traffic,
                     "traffic_over_algorithmic": traffic_ratio,
                     "kernel": kname,
                     "launches": launches, "avg_launch_ms": 1e3 * avg_s,
                     "avg_bytes_per_launch": avg_bytes,
                     "avg_flop_per_launch": flops_per_launch,
                     "hbm_gbs": avg_bytes / avg_s / 1e9 if avg_s > 0 else 0.0},
        "cpu_baseline": cpu,
        "extra": extra,
        "iters_mean_per_size": {str(n): float(np.mean(iters_mixed[c])) for c, n in enumerate(sizes)},
        "outer_mean_per_size": {str(n): float(np.mean(outer_mixed[c])) for c, n in enumerate(sizes)},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
