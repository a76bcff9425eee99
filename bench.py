#!/usr/bin/env python3
"""Monte Carlo solves/sec of the Hexa8 cantilever on one MI355X.

Workload (one "step"): the batched mixed-precision Monte Carlo job of
1024 samples whose cube-mesh sizes are drawn uniformly from
n in {16, 24, 32, 40, 48} elements per side; per sample a lognormal
Karhunen-Loeve modulus field (40 terms, exponential covariance, correlation
length 0.25, mean 2.1e5, COV 0.3) evaluated at element centroids, FP64
element integration rounded once to FP32 stored values, FP32 PCG under FP64
residual replacement to |f - K u|/|f| < 1e-9.  Inputs (meshes, KL tables,
KL coefficients, loads) are resident on the device before the timed region;
each step re-evaluates the modulus, re-assembles and re-solves every sample.
The PCG preconditioner is a geometric multigrid V-cycle per sample
(--precond mg, default) or Jacobi (--precond jacobi, the reference's
SolverPCG choice); the other one is timed beside it for comparison.

Reported beside the headline (all in "extra"):
  - each class's first sample against the CPU oracle's own solve (CSR
    Jacobi-PCG to 1e-12): the independent error of the headline;
  - the same job in plain FP64 (batched) and its error against a tightly
    converged FP64 batch of all 1024 samples;
  - the single-sample FP64 GPU Jacobi-PCG per mesh size;
  - 256 samples of 32^3 in FP64 (Jacobi-PCG to 1e-10);
  - one 64^3 FP64 Jacobi-PCG solve to 1e-10;
  - the batched Newmark dynamics of a 48x48x24 block (64 samples, 500 steps);
  - the roofline of the kernels that take the most time, from HIP events on
    the context stream;
  - the CPU baseline: the oracle's restatement of the reference per-sample
    loop, one sample per mesh size, pinned to one core, and the same on
    every core of the host share.
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

METRIC = "MC solves/sec at 1e-8 rel. err, batched mixed vs plain FP64, 1 MI355X"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3      # MI355X published FP32 rate (packed / MFMA)
FP32_VALU_NOPACK_TFLOPS = 78.65  # the same without packed FP32 (disabled in this build)
HOST_CORE_SHARE = 16          # host cores one GPU gets on the benchmark box


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
    p.add_argument("--oracle-tol", type=float, default=1e-12,
                   help="tolerance of the CPU oracle solves the headline is checked against")
    p.add_argument("--fp64-tol", type=float, default=1e-13,
                   help="tolerance of the tightly converged FP64 GPU batch")
    p.add_argument("--max-it", type=int, default=20000)
    p.add_argument("--max-outer", type=int, default=30)
    p.add_argument("--graph-iters", type=int, default=12)
    p.add_argument("--tile", type=int, default=0,
                   help="lane tile of the batches (0: tile and slot unroll from the autotuner)")
    p.add_argument("--unroll", type=int, default=3, help="slot unroll when --tile is given")
    p.add_argument("--tune-samples", type=int, default=64,
                   help="samples per class of the autotuner's trial batches")
    p.add_argument("--period", type=int, default=None,
                   help="replacement check period (iterations; default 2 with multigrid, 8 with Jacobi)")
    p.add_argument("--drop", type=float, default=None,
                   help="FP32 residual drop that triggers a replacement (default 1e-4 with "
                        "multigrid, 0.1 with Jacobi)")
    p.add_argument("--refine", choices=["reliable", "restart"], default="reliable")
    p.add_argument("--matrix-cores", type=int, default=0,
                   help="1: box-grid smoothing products on the matrix cores, 0: on the vector ALU "
                        "(the other form is timed beside the headline)")
    p.add_argument("--lanczos-steps", type=int, default=None,
                   help="Lanczos steps of the multigrid smoothing interval (0: Gershgorin bound; "
                        "default: the library's)")
    p.add_argument("--lanczos-margin", type=float, default=None,
                   help="margin over the top Ritz value (default: the library's)")
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
                   help="threads of the multi-core CPU baseline (-1: every core of the host "
                        "share, 0: skip it)")
    p.add_argument("--no-single", action="store_true", help="skip the single-sample FP64 runs")
    p.add_argument("--no-c2", action="store_true", help="skip the 64^3 FP64 single solve")
    p.add_argument("--no-c3", action="store_true", help="skip the 256 x 32^3 FP64 batch")
    p.add_argument("--no-c5", action="store_true", help="skip the batched Newmark dynamics")
    p.add_argument("--no-persistent", action="store_true",
                   help="skip the cooperative persistent-launch runs of the C2/C3 legs")
    p.add_argument("--only", choices=["c2", "c3", "c5"], default=None,
                   help="run only this comparison workload (profiling)")
    p.add_argument("--newmark", type=int, default=0, metavar="STEPS",
                   help="instead of the Monte Carlo job, time STEPS Newmark steps of the "
                        "batched dynamics workload (48x48x24 block, 64 samples)")
    p.add_argument("--newmark-steps", type=int, default=500,
                   help="time steps of the dynamics workload in the default run")
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
    replacement waits for a 1e-4 drop; Jacobi-PCG keeps 8 and 0.1."""
    mg = precond == "mg"
    return (args.period if args.period is not None else (2 if mg else 8),
            args.drop if args.drop is not None else (1e-4 if mg else 0.1))


def make_batch(msa, ctx, meshes, models, counts, precision, tile, table, xi, args, precond=None,
               cov=None):
    b = msa.Batch(ctx, meshes, counts, precision, tile)
    precond = precond or args.precond
    if precond == "mg":
        b.set_preconditioner(msa.MSA_PRECOND_MULTIGRID)
        if args.lanczos_steps is not None or args.lanczos_margin is not None:
            k0, m0 = b.smoothing_interval()
            b.set_smoothing_interval(k0 if args.lanczos_steps is None else args.lanczos_steps,
                                     m0 if args.lanczos_margin is None else args.lanczos_margin)
    if precision == msa.MSA_MIXED:
        period, drop = refinement(args, precond)
        b.set_refinement(msa.MSA_REFINE_RELIABLE if args.refine == "reliable"
                         else msa.MSA_REFINE_RESTART, period, drop)
    off = 0
    for c, m in enumerate(models):
        b.set_kl(c, table, m.domain, 2.1e5, args.cov if cov is None else cov)
        b.set_xi(c, xi[off:off + counts[c]])
        b.set_rhs(c, m.f)
        off += counts[c]
    ctx.synchronize()
    return b


def host_info():
    """Host of this run: core counts, CPU model and clocks (from /proc and
    /sys, no external tools)."""
    aff = sorted(os.sched_getaffinity(0))
    model, mhz = None, []
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name") and model is None:
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("cpu MHz"):
                    mhz.append(float(line.split(":", 1)[1]))
    except OSError:
        pass
    fmax = None
    try:
        with open(f"/sys/devices/system/cpu/cpu{aff[0]}/cpufreq/cpuinfo_max_freq") as fh:
            fmax = int(fh.read().strip()) / 1000.0
    except (OSError, ValueError, IndexError):
        pass
    return dict(nproc=os.cpu_count(), affinity_cpus=len(aff), model=model,
                mhz_mean=float(np.mean(mhz)) if mhz else None, mhz_max=fmax)


# ------------------------------------------------------------------ work models
def mf_work(b, models):
    """FP32 operations and algorithmic HBM bytes of all finest-level
    matrix-free smoothing launches of one solve (None unless every class
    with a hierarchy uses them).  Per cycle and active sample four products;
    each does, for every (free node, incident cell) pair, 72 FMA (three
    rows of the unit element matrix times the cell's 24 values) and 3 FMA
    (modulus scaling): 150 flop.  Bytes per product and sample: the moduli
    (8 per element) and 4 vectors of 4*3*nb on average over the four steps;
    the grid tables once per launch.  The set-up's Lanczos steps (one RESID
    product each, every sample) count as products too."""
    flops = byts = 0.0
    lz_steps = b.smoothing_interval()[0]
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
        prods = cyc.sum() * 4 + lz_steps * len(cyc)
        flops += prods * inc * 150.0
        byts += prods * (8.0 * conn.shape[0] + 4 * 4 * 3 * nb)
        byts += (float(cyc.max()) * 4 + lz_steps) * 4.0 * (len(free_node) + conn.shape[0] + nb)
    return flops, byts


def smooth_bytes(b, models, rs):
    """Algorithmic bytes of all finest-level stored-value smoothing launches
    of one solve: per cycle and active sample the four steps read the values
    four times and move 16 vectors, plus the shared block-column indices and
    slice pointers once per launch."""
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


def kernel_a_bytes(b, rs, full=False):
    """Algorithmic bytes moved by all SpMV+dot (k_pcg_a) launches of one
    solve: per active sample and iteration the matrix values and the four
    vectors (z, p_old read; p_new, q written); the shared block-column
    indices and slice pointers once per launch of a class with active
    samples.  Classes whose product reads the upper triangle (box grids,
    one-sample tiles) count its values, diagonal included, and no indices;
    full=True prices every class at the full-storage bytes."""
    total = 0.0
    for c, mesh in enumerate(b.meshes):
        st = b.stats(c)
        nnz, n = mesh.info["nnz"], 3 * mesh.info["blocks"]
        its = st["iters"].astype(np.float64)
        if b.layout_info(c)["upper_triangle_product"] and not full:
            total += its.sum() * (rs * 0.5 * (nnz + n) + 4 * rs * n)
            continue
        total += its.sum() * (rs * nnz + 4 * rs * n)
        idx = 4.0 * mesh.info["block_nnz"] + 4.0 * (mesh.info["blocks"] + 1)
        total += float(its.max()) * idx
    return total


def mf_pcg_work(b, models):
    """FP32 operations and algorithmic HBM bytes of all matrix-free PCG
    product launches (k_pcg_p + k_pcg_a_mf) of one solve, None unless
    every class takes that product.  Per iteration and active sample: the
    same 150 flop per (free node, incident cell) pair as the matrix-free
    smoothing products; bytes: z and p_old read and p written by the p
    pass, p gathered and q written by the product (5 vectors of 4*3*nb),
    and the moduli (8 per element)."""
    flops = byts = 0.0
    for c, m in enumerate(models):
        if b.product_form(c) != "matrix_free":
            return None
        its = b.stats(c)["iters"].astype(np.float64).sum()
        conn = np.asarray(m.conn).reshape(-1, 8)
        free_node = (np.asarray(m.dof).reshape(-1, 3) >= 0).any(axis=1)
        inc = float(free_node[conn].sum())
        nb = float(free_node.sum())
        flops += its * inc * 150.0
        byts += its * (5 * 4 * 3 * nb + 8.0 * conn.shape[0])
    return flops, byts


def counters(name):
    """Measured/algorithmic HBM bytes of a kernel from its rocprofv3
    FETCH_SIZE/WRITE_SIZE passes (profiles/<name>), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    cj = json.load(open(path))
    return (cj["fetch_bytes"] + cj["write_bytes"]) / (cj["algorithmic_read"] +
                                                     cj["algorithmic_write"])


def back_to_back(roof, b, reps=20, rs=8):
    """Adds the SpMV+dot kernel timed over `reps` back-to-back launches with
    every sample active (no per-launch event overhead) to a roofline entry
    and makes it the entry's achieved figure; the per-launch HIP-event
    figure of the solve stays beside it.  Overwrites the batch's PCG
    iterates."""
    ms = b.time_spmv(reps)
    byts = b.spmv_bytes()["kernel"]
    roof["per_launch_events"] = dict(achieved=roof["achieved"], avg_launch_ms=roof["avg_launch_ms"],
                                     avg_bytes_per_launch=roof["avg_bytes_per_launch"])
    roof["achieved"] = byts / (ms / 1e3) / 1e9
    roof["frac"] = roof["achieved"] / HBM_PEAK_GBS
    roof["avg_launch_ms"] = ms
    roof["avg_bytes_per_launch"] = byts
    roof["timing"] = f"{reps} back-to-back launches, every sample active, HIP events"
    if roof.get("traffic_over_algorithmic"):
        roof["traffic"] = byts * roof["traffic_over_algorithmic"]
    if "full_storage_equivalent" in roof:
        full = 0.0
        for c, mesh in enumerate(b.meshes):
            nnz, n = mesh.info["nnz"], 3 * mesh.info["blocks"]
            full += b.n_samples[c] * (rs * nnz + 4 * rs * n)
            full += 4.0 * mesh.info["block_nnz"] + 4.0 * (mesh.info["blocks"] + 1)
        roof["full_storage_equivalent"].update(gbs=full / (ms / 1e3) / 1e9,
                                               avg_bytes_per_launch=full)
    return roof


def spmv_name(b, real, tile):
    """Name of the SpMV+dot kernel entry, saying which product it runs."""
    if all(b.product_form(c) == "matrix_free" for c in range(len(b.meshes))):
        return ("k_pcg_p + k_pcg_a_mf (p = z + beta p_old, then the matrix-free FP32 PCG "
                "product K p on box grids from the cell moduli, the corners' differences from "
                "the node's value, + p.q; time of both kernels)")
    if all(b.layout_info(c)["upper_triangle_product"] for c in range(len(b.meshes))):
        return (f"k_pcg_a<{real},{tile}> (symmetric SpMV from the stored upper triangle "
                "+ p update + p.q)")
    return f"k_pcg_a<{real},{tile}> (SpMV + p update + p.q)"


def spmv_roofline(b, kt, rs, name, counter_file, models=None):
    """Roofline entry of k_pcg_a from one kernel-timing solve (the
    matrix-free product: FP32 VALU, like the smoothing products)."""
    launches = kt["spmv_dot"]["launches"]
    a_s = kt["spmv_dot"]["ms"] / 1e3 / max(launches, 1)
    mfw = mf_pcg_work(b, models) if models is not None else None
    if mfw is not None:
        flop, byts = mfw[0] / max(launches, 1), mfw[1] / max(launches, 1)
        ach = flop / a_s / 1e12 if a_s > 0 else 0.0
        ratio = counters("k_pcg_a_mf_counters.json")
        return dict(bound="valu", achieved=ach, peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=ach / FP32_PEAK_TFLOPS,
                    frac_of_valu_without_packed_fp32=ach / FP32_VALU_NOPACK_TFLOPS,
                    traffic=byts * ratio if ratio is not None else None,
                    traffic_over_algorithmic=ratio,
                    kernel=name, launches=launches, avg_launch_ms=1e3 * a_s,
                    avg_flop_per_launch=flop, avg_bytes_per_launch=byts,
                    hbm_gbs=byts / a_s / 1e9 if a_s > 0 else 0.0,
                    ms_total=kt["spmv_dot"]["ms"])
    a_bytes = kernel_a_bytes(b, rs) / max(launches, 1)
    ratio = counters(counter_file)
    ach = a_bytes / a_s / 1e9 if a_s > 0 else 0.0
    out = dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
               frac=ach / HBM_PEAK_GBS,
               traffic=a_bytes * ratio if ratio is not None else None,
               traffic_over_algorithmic=ratio, kernel=name, launches=launches,
               avg_launch_ms=1e3 * a_s, avg_bytes_per_launch=a_bytes,
               ms_total=kt["spmv_dot"]["ms"])
    full = kernel_a_bytes(b, rs, full=True) / max(launches, 1)
    if full != a_bytes and a_s > 0:
        # the symmetric product: the same launches priced at the bytes a
        # full-storage product moves (both triangles and the indices), to
        # compare with the full-storage kernel; `achieved` stays the
        # product's own bytes
        out["full_storage_equivalent"] = dict(
            gbs=full / a_s / 1e9, avg_bytes_per_launch=full,
            note="the same launches priced at full-storage bytes (both triangles + indices), "
                 "to compare with a full-storage SpMV; not the kernel's own traffic and not "
                 "a fraction of the HBM peak")
    return out


def smoother_roofline(bm, models, kt, args, matrix_cores=True):
    """Roofline entry of the finest level's smoothing products (matrix-core
    or vector-ALU element products: the algorithmic flops are the same)."""
    launches = kt["multigrid"]["launches"]
    if launches == 0:
        return None
    avg_s = kt["multigrid"]["ms"] / 1e3 / launches
    mf = mf_work(bm, models)
    if mf is not None:
        flop = mf[0] / launches
        ach = flop / avg_s / 1e12 if avg_s > 0 else 0.0
        ratio = counters("k_mg_smooth_mf_counters.json")
        byts = mf[1] / launches
        if matrix_cores:
            bound, name = "mfma", ("k_mg_smooth_mfma<STEP> on the finest level: matrix-free FP32 "
                                   "products on box grids (PRE, RESID, POST1, POST2), element "
                                   "products by v_mfma_f32_16x16x4f32")
        else:
            bound, name = "valu", ("k_mg_smooth<float,1,STEP,2> on the finest level: matrix-free "
                                   "FP32 products on box grids (PRE, RESID, POST1, POST2), FP32 VALU")
        return dict(bound=bound, achieved=ach, peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=ach / FP32_PEAK_TFLOPS,
                    frac_of_valu_without_packed_fp32=ach / FP32_VALU_NOPACK_TFLOPS,
                    traffic=byts * ratio if ratio is not None else None,
                    traffic_over_algorithmic=ratio,
                    kernel=name,
                    launches=launches, avg_launch_ms=1e3 * avg_s, avg_flop_per_launch=flop,
                    avg_bytes_per_launch=byts, hbm_gbs=byts / avg_s / 1e9,
                    ms_total=kt["multigrid"]["ms"])
    byts = smooth_bytes(bm, models, 4) / launches
    ratio = counters("k_mg_smooth_counters.json")
    ach = byts / avg_s / 1e9
    return dict(bound="hbm", achieved=ach, peak=HBM_PEAK_GBS, unit="GB/s",
                frac=ach / HBM_PEAK_GBS, traffic=byts * ratio if ratio is not None else None,
                traffic_over_algorithmic=ratio,
                kernel=f"k_mg_smooth<float,{args.tile},STEP> on the finest level (stored values)",
                launches=launches, avg_launch_ms=1e3 * avg_s, avg_bytes_per_launch=byts,
                ms_total=kt["multigrid"]["ms"])


# ------------------------------------------------------------------ oracle checks
def oracle_true_residual(O, model, E, u, nu=0.3):
    """|f - K u|/|f| with the oracle's own element matrices and CSR assembly
    (independent of the GPU's operator)."""
    ke, _ = O.ke_mesh(model.conn, model.xyz, E, nu)
    rp, col, val = O.csr_assemble(model.conn, model.dof, model.n_free, ke)
    r = model.f - O.csr_multiply(rp, col, val, u)
    return float(np.linalg.norm(r) / np.linalg.norm(model.f))


# ------------------------------------------------------------------ legs
def persistent_run(b, ctx, kw, ref, args):
    """The same solve as one cooperative persistent launch: wall time of one
    run (after one warm-up) and whether it gives the graph path's bits
    (None with --no-persistent)."""
    if args.no_persistent:
        return None
    b.set_persistent(True)
    b.run(**kw)
    ctx.synchronize()
    t = time.perf_counter()
    b.run(**kw)
    ctx.synchronize()
    el = time.perf_counter() - t
    same = all(np.array_equal(b.solution(0, s), u) for s, u in ref.items())
    info = b.persistent_info()
    b.set_persistent(False)
    return dict(s_per_solve=el, same_bits_as_graph=bool(same), workgroups=info["workgroups"],
                launches=info["launches"])


def leg_c3(msa, ctx, args):
    """256 samples of the 32^3 cantilever, lognormal KL modulus (COV 0.2),
    FP64 element integration, assembly and Jacobi-PCG to 1e-10, one batch."""
    from msolve_amd.stochastic import kl_table, standard_normal
    from oracle import oracle as O
    ns, n = 256, 32
    model = msa.cantilever(n)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    table = kl_table([1.0, 1.0, 1.0], args.corr, args.terms)
    xi = standard_normal(args.seed + 3, ns, args.terms)
    b = make_batch(msa, ctx, [mesh], [model], [ns], msa.MSA_FP64, 1, table, xi, args,
                   precond="jacobi", cov=0.2)
    kw = dict(nu=0.3, tol=1e-10, max_it=args.max_it, graph_iters=args.graph_iters)
    b.run(**kw)
    ctx.synchronize()
    t = time.perf_counter()
    b.run(**kw)
    ctx.synchronize()
    el = time.perf_counter() - t
    st = b.stats(0)
    b.set_kernel_timing(True)
    b.run(**kw)
    kt = b.kernel_timing()
    b.set_kernel_timing(False)
    roof = spmv_roofline(b, kt, 8, spmv_name(b, "double", 1) + ", 256 x 32^3",
                         "k_pcg_a_fp64_counters.json")
    u = {s: b.solution(0, s) for s in (0, ns - 1)}
    pers = persistent_run(b, ctx, kw, u, args)
    back_to_back(roof, b)
    # independent check of two samples: the oracle's own CSR operator
    otab = O.kl_table([1.0, 1.0, 1.0], args.corr, args.terms)
    res = {}
    for s in (0, ns - 1):
        E, bad = O.kl_field(model.conn, model.xyz, otab, xi[s], model.domain, 2.1e5, 0.2)
        res[str(s)] = oracle_true_residual(O, model, E, u[s])
    b.close()
    mesh.close()
    out = dict(workload="256 lognormal-KL samples (COV 0.2) of the 32^3 cantilever, FP64 "
                        "Jacobi-PCG to |f-Ku|/|f| < 1e-10, one batch",
               solves_per_s=ns / el, s_per_batch=el, iters_mean=float(st["iters"].mean()),
               iters_max=int(st["iters"].max()),
               all_ok=bool((st["status"] == msa.MSA_OK).all()),
               oracle_true_residual=res, roofline=roof, persistent_launch=pers,
               kernel_timing_ms={k: v["ms"] for k, v in kt.items()})
    log(f"[bench] C3 256x32^3 FP64: {el:.2f} s -> {ns / el:.1f} solves/s, "
        f"k_pcg_a<double> {roof['achieved']:.0f} GB/s, residuals {res}" +
        (f"; persistent launch {pers['s_per_solve']:.2f} s (same bits: "
         f"{pers['same_bits_as_graph']})" if pers else ""))
    return out


def leg_c2(msa, ctx, args):
    """One FP64 Jacobi-PCG solve of the 64^3 cantilever (E = 2.1e5) to 1e-10."""
    from oracle import oracle as O
    n = 64
    model = msa.cantilever(n)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    b = msa.Batch(ctx, [mesh], [1], msa.MSA_FP64, 1)
    E = np.full(model.n_elem, 2.1e5)
    b.set_modulus(0, E)
    b.set_rhs(0, model.f)
    kw = dict(nu=0.3, tol=1e-10, max_it=args.max_it, graph_iters=args.graph_iters)
    b.run(**kw)
    ctx.synchronize()
    t = time.perf_counter()
    b.run(**kw)
    ctx.synchronize()
    el = time.perf_counter() - t
    st = b.stats(0)
    b.set_kernel_timing(True)
    b.run(**kw)
    kt = b.kernel_timing()
    b.set_kernel_timing(False)
    roof = spmv_roofline(b, kt, 8, spmv_name(b, "double", 1) + ", one 64^3 sample",
                         "k_pcg_a_fp64_counters.json")
    u = b.solution(0, 0)
    pers = persistent_run(b, ctx, kw, {0: u}, args)
    back_to_back(roof, b)
    res = oracle_true_residual(O, model, E, u)
    b.close()
    mesh.close()
    out = dict(workload="64^3 cantilever (811200 DOFs), E 2.1e5, FP64 Jacobi-PCG to 1e-10",
               s_per_solve=el, iters=int(st["iters"][0]), status=int(st["status"][0]),
               ms_per_iteration=1e3 * el / max(int(st["iters"][0]), 1),
               oracle_true_residual=res, roofline=roof, persistent_launch=pers,
               kernel_timing_ms={k: v["ms"] for k, v in kt.items()})
    log(f"[bench] C2 64^3 FP64: {el:.3f} s, {int(st['iters'][0])} its, k_pcg_a<double> "
        f"{roof['achieved']:.0f} GB/s, oracle residual {res:.2e}" +
        (f"; persistent launch {pers['s_per_solve']:.3f} s (same bits: "
         f"{pers['same_bits_as_graph']})" if pers else ""))
    return out


def leg_newmark(msa, ctx, args, nsteps):
    """Batched Newmark dynamics: 48x48x24 Hexa8 block (1 x 1 x 0.5, clamped
    at x = 0), consistent mass, density 7.85e-9, Newmark beta 0.25 / gamma
    0.5, dt 1e-4, base acceleration 1.0 sin(2 pi 20 t) in Z, lognormal KL
    modulus (COV 0.2), mixed-precision PCG warm-started from the previous
    displacement every step.  Rate = samples x steps / s."""
    from msolve_amd.stochastic import kl_table, standard_normal
    ns = args.newmark_samples
    model = msa.cantilever(48, 48, 24, lz=0.5)
    mesh = msa.Mesh(ctx, model.xyz, model.conn, model.dof, model.n_free)
    table = kl_table(list(model.domain), args.corr, args.terms)
    xi = standard_normal(args.seed, ns, args.terms)
    acc = msa.sinusoidal_history(0.05, 1e-4, 0.05, 1.0, nsteps)

    def make(precision, nsamp):
        b = msa.Batch(ctx, [mesh], [nsamp], precision, 1)
        if args.precond == "mg":
            b.set_preconditioner(msa.MSA_PRECOND_MULTIGRID)
            if args.lanczos_steps is not None or args.lanczos_margin is not None:
                k0, m0 = b.smoothing_interval()
                b.set_smoothing_interval(k0 if args.lanczos_steps is None else args.lanczos_steps,
                                         m0 if args.lanczos_margin is None else args.lanczos_margin)
        if precision == msa.MSA_MIXED:
            period, drop = refinement(args, args.precond)
            b.set_refinement(msa.MSA_REFINE_RELIABLE, period, drop)
        b.set_kl(0, table, model.domain, 2.1e5, 0.2)
        b.set_xi(0, xi[:nsamp])
        b.eval_kl()
        b.set_dynamics(7.85e-9, 0.0, 0.0, 0.25, 0.5, 1e-4)
        b.set_probe(0, model.n_free - 1)
        b.assemble(0.3)
        ctx.synchronize()
        return b

    b = make(msa.MSA_MIXED, ns)
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
    nchk = min(4, ns)
    b64 = make(msa.MSA_FP64, nchk)
    b64.newmark(acc, 2, 1e-12, args.max_it, args.graph_iters)
    errs = [float(np.linalg.norm(hist[s] - b64.history(0, s)) / np.linalg.norm(b64.history(0, s)))
            for s in range(nchk)]
    b64.close()
    mesh.close()
    log(f"[bench] Newmark: {nsteps} steps x {ns} samples in {el:.2f} s -> "
        f"{ns * nsteps / el:.1f} sample-steps/s, probe error vs FP64 {max(errs):.2e}")
    return dict(
        workload=f"batched Newmark: {ns} lognormal-KL samples (COV 0.2) of a 48x48x24 Hexa8 "
                 f"block, rho 7.85e-9, beta 0.25, gamma 0.5, dt 1e-4, 20 Hz base acceleration, "
                 f"{'multigrid' if args.precond == 'mg' else 'Jacobi'}-preconditioned mixed PCG "
                 f"to {args.newmark_tol} warm-started from the previous displacement",
        sample_steps_per_s=ns * nsteps / el, samples=ns, steps=nsteps,
        ms_per_time_step=1e3 * el / nsteps, n_free=model.n_free,
        all_ok=bool((st["status"] == msa.MSA_OK).all()),
        iters_per_step_mean=float(np.mean(its)) / nsteps,
        probe_history_rel_err_vs_fp64=max(errs), **extra)


def bench_newmark(args):
    import msolve_amd as msa
    ctx = msa.Context(0)
    r = leg_newmark(msa, ctx, args, args.newmark)
    ctx.close()
    out = {
        "metric": "Newmark sample-steps/sec, batched mixed warm-started PCG, 1 MI355X",
        "value": r["sample_steps_per_s"], "unit": "sample-steps/s", "n_gpus": 1,
        "steps": args.newmark, "ms_per_step": r["ms_per_time_step"], "higher_is_better": True,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": r["workload"], "precond": args.precond},
        "extra": r,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(args, sizes, counts, xi, u_mixed):
    """The oracle's CSR/Jacobi restatement of the reference per-sample loop
    (KL modulus, Ke, assembly, PCG to the headline's tolerance), one sample
    per mesh size: pinned to one core, then on every core of the host
    share.  Also the oracle's own tightly converged solve of each class's
    first sample, against which the headline's displacements are compared."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O
    otab = O.kl_table([1.0, 1.0, 1.0], args.corr, args.terms)
    host = host_info()
    aff = sorted(os.sched_getaffinity(0))
    core = aff[0]
    per, u_base, t_meas = [], [], 0.0
    off = 0
    os.sched_setaffinity(0, {core})
    try:
        for c, n in enumerate(sizes):
            out = O.mc_reference(n, otab, xi[off:off + 1], cov=args.cov, tol=args.tol,
                                 keep_u=True, storage="csr")
            per.append(out["times"]["all"])
            u_base.append(out["u"][0])
            t_meas += out["times"]["all"]
            off += counts[c]
            log(f"[bench] cpu {n}^3 (core {core}): {out['times']['all']:.2f} s, "
                f"{int(out['iters'][0])} its")
    finally:
        os.sched_setaffinity(0, set(aff))
    N = sum(counts)
    t_cpu = sum(counts[c] * per[c] for c in range(len(sizes)))
    cpu = dict(value=N / t_cpu, unit="solves/s", cores=1, kind="port",
               sample=f"one sample per mesh size {sizes} (first of each class), oracle "
                      f"CSR/Jacobi-PCG FP64 restatement of the reference per-sample loop "
                      f"(KL modulus, Ke, assembly, PCG to {args.tol}), serial, pinned to host "
                      f"core {core}; rate = {N} / sum(class count x measured time), "
                      f"{t_meas:.1f} s measured",
               host=host, per_size_s=dict(zip(map(str, sizes), per)))
    # the oracle's tight solves (one thread per class, in parallel)
    offs = np.cumsum([0] + counts[:-1]).tolist()

    def tight(c):
        return O.mc_reference(sizes[c], otab, xi[offs[c]:offs[c] + 1], cov=args.cov,
                              tol=args.oracle_tol, keep_u=True, storage="csr")["u"][0]

    t0 = time.perf_counter()
    with ThreadPoolExecutor(len(sizes)) as ex:
        u_ref = list(ex.map(tight, range(len(sizes))))
    log(f"[bench] oracle solves to {args.oracle_tol}: {time.perf_counter() - t0:.1f} s wall")
    rel = lambda a, r: float(np.linalg.norm(a - r) / np.linalg.norm(r))
    indep = {str(n): rel(u_mixed[c], u_ref[c]) for c, n in enumerate(sizes)}
    cpu_err = {str(n): rel(u_base[c], u_ref[c]) for c, n in enumerate(sizes)}
    cpu["rel_err_vs_oracle_tight"] = cpu_err
    parallel = None
    nthr = args.cpu_threads
    if nthr < 0:
        nthr = min(len(aff), HOST_CORE_SHARE)
    if nthr > 1:
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
        parallel = dict(
            value=rate, unit="solves/s", cores=nthr, kind="port",
            sample=f"{nthr} threads (every core of the host share one GPU gets; the host shows "
                   f"nproc {host['nproc']}, {host['affinity_cpus']} usable), each one further "
                   f"sample per mesh size {sizes}, the same oracle restatement; rate = {nthr} x "
                   f"{N} / sum(class count x mean per-size time under load), {wall:.1f} s wall")
        log(f"[bench] cpu x{nthr}: {rate:.2f} solves/s ({wall:.1f} s wall)")
    return cpu, parallel, indep


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the GPU box has one, and the "
                                   "sharding of the Monte Carlo samples over several GPUs is a "
                                   "design, not a measured path",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.newmark:
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
    import msolve_amd as msa
    if args.only:
        ctx = msa.Context(0)
        fn = {"c2": leg_c2, "c3": leg_c3}.get(args.only)
        r = fn(msa, ctx, args) if fn else leg_newmark(msa, ctx, args, args.newmark_steps)
        ctx.close()
        print(json.dumps({"only": args.only, "result": r}), flush=True)
        return
    if args.only_mixed:
        args.no_fp64 = args.no_cpu = args.no_single = True
        args.no_c2 = args.no_c3 = args.no_c5 = True
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

    # ------------------------------------------------ schedule: autotuner
    tuned = None
    if args.tile <= 0:
        t = time.perf_counter()
        pk = msa.MSA_PRECOND_MULTIGRID if args.precond == "mg" else msa.MSA_PRECOND_JACOBI
        tuned = msa.autotune_schedule(ctx, meshes, args.tune_samples, msa.MSA_MIXED, pk)
        args.tile, args.unroll = tuned["tile"], tuned["unroll"]
        log(f"[bench] autotuner ({time.perf_counter() - t:.1f} s): tile {args.tile}, unroll "
            f"{args.unroll}; ms per sample {tuned['table']}")

    # ------------------------------------------------ headline: mixed, batched
    bm = make_batch(msa, ctx, meshes, models, counts, msa.MSA_MIXED, args.tile, table, xi, args)
    bm.set_schedule(args.unroll)
    bm.set_matrix_cores(bool(args.matrix_cores))
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

    # rooflines: one more step with HIP events around every launch on the
    # context stream; the kernel with the most time per step is `roofline`
    bm.set_kernel_timing(True)
    bm.run(**run_kw)
    kt = bm.kernel_timing()
    bm.set_kernel_timing(False)
    log(f"[bench] kernel timing: {json.dumps(kt)}")
    roof_a = spmv_roofline(bm, kt, 4, spmv_name(bm, "float", args.tile), "k_pcg_a_counters.json",
                           models)
    roof_s = (smoother_roofline(bm, models, kt, args, bool(args.matrix_cores))
              if args.precond == "mg" else None)
    if roof_s is not None and roof_s["ms_total"] > roof_a["ms_total"]:
        roofline, secondary = roof_s, roof_a
    else:
        roofline, secondary = roof_a, roof_s
    u_mixed = [[bm.solution(c, s) for s in range(counts[c])] for c in range(len(models))]
    iters_mixed = [s["iters"].tolist() for s in stats]
    outer_mixed = [s["outer"].tolist() for s in stats]
    full_storage = stored = None
    mf_all = all(bm.product_form(c) == "matrix_free" for c in range(len(models)))
    sym_all = all(bm.layout_info(c)["upper_triangle_product"] for c in range(len(models)))
    if not args.only_mixed and mf_all:
        # the same batch with the PCG product from the stored upper triangle
        # (msa_batch_set_matrix_free_product(b, 0)): its HBM roofline and
        # the solve rate beside the matrix-free product's
        bm.set_matrix_free_product(False)
        bm.run(**run_kw)
        ctx.synchronize()
        t = time.perf_counter()
        bm.run(**run_kw)
        ctx.synchronize()
        ts_ = time.perf_counter() - t
        bm.set_kernel_timing(True)
        bm.run(**run_kw)
        kts = bm.kernel_timing()
        bm.set_kernel_timing(False)
        stored = dict(
            solves_per_s=N / ts_,
            iters_mean_per_size={str(n): float(np.mean(bm.stats(c)["iters"]))
                                 for c, n in enumerate(sizes)},
            roofline=spmv_roofline(bm, kts, 4, spmv_name(bm, "float", args.tile),
                                   "k_pcg_a_counters.json"),
            kernel_timing_ms_per_step={k: v["ms"] for k, v in kts.items()})
        log(f"[bench] mixed with the stored upper-triangle product: {ts_:.2f} s -> "
            f"{N / ts_:.1f} solves/s, k_pcg_a {stored['roofline']['achieved']:.0f} GB/s "
            f"({stored['roofline']['frac']:.3f} of peak)")
    if not args.only_mixed and sym_all:
        # the same batch with the full image (both triangles stored and
        # streamed, msa_batch_set_symmetric_product(b, 0)): the plain SpMV's
        # rate beside the symmetric product's
        bm.set_symmetric_product(False)
        bm.run(**run_kw)
        ctx.synchronize()
        t = time.perf_counter()
        bm.run(**run_kw)
        ctx.synchronize()
        tf = time.perf_counter() - t
        bm.set_kernel_timing(True)
        bm.run(**run_kw)
        ktf = bm.kernel_timing()
        bm.set_kernel_timing(False)
        full_storage = dict(
            solves_per_s=N / tf,
            roofline=spmv_roofline(bm, ktf, 4, spmv_name(bm, "float", args.tile),
                                   "k_pcg_a_full_counters.json"),
            kernel_timing_ms_per_step={k: v["ms"] for k, v in ktf.items()})
        log(f"[bench] mixed with the full-storage SpMV: {tf:.2f} s -> {N / tf:.1f} solves/s, "
            f"k_pcg_a {full_storage['roofline']['achieved']:.0f} GB/s "
            f"({full_storage['roofline']['frac']:.3f} of peak)")
        bm.set_symmetric_product(True)
    bm.set_matrix_free_product(True)
    other_form = None
    if args.precond == "mg" and not args.only_mixed:
        # the same batch with the smoothing products in the other form
        mc = not bool(args.matrix_cores)
        bm.set_matrix_cores(mc)
        bm.run(**run_kw)
        ctx.synchronize()
        t = time.perf_counter()
        bm.run(**run_kw)
        ctx.synchronize()
        tv = time.perf_counter() - t
        bm.set_kernel_timing(True)
        bm.run(**run_kw)
        ktv = bm.kernel_timing()
        bm.set_kernel_timing(False)
        other_form = dict(matrix_cores=mc, solves_per_s=N / tv,
                          roofline=smoother_roofline(bm, models, ktv, args, mc),
                          kernel_timing_ms_per_step={k: v["ms"] for k, v in ktv.items()})
        log(f"[bench] mixed with the {'matrix-core' if mc else 'vector-ALU'} smoother: "
            f"{tv:.2f} s -> {N / tv:.1f} solves/s")
        bm.set_matrix_cores(bool(args.matrix_cores))
    bm.close()

    extra = {"roofline_second": secondary, "mixed_other_smoother_form": other_form,
             "mixed_stored_upper_triangle_product": stored,
             "mixed_full_storage_spmv": full_storage,
             "kernel_timing_ms_per_step": {k: v["ms"] for k, v in kt.items()},
             "kernel_launches_per_step": {k: v["launches"] for k, v in kt.items()}}
    # ------------------------------------------------ the other preconditioner
    other = "jacobi" if args.precond == "mg" else "mg"
    if not args.no_compare and not args.only_mixed:
        bo = make_batch(msa, ctx, meshes, models, counts, msa.MSA_MIXED, args.tile, table, xi,
                        args, precond=other)
        bo.set_schedule(args.unroll)
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
    # ------------------------------------------------ FP64 batched
    max_err = None
    if not args.no_fp64:
        b64 = make_batch(msa, ctx, meshes, models, counts, msa.MSA_FP64, args.tile, table, xi, args)
        b64.run(**run_kw)
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
        b64.close()
        # every sample against a tightly converged FP64 batch (a batch of its
        # own, one captured graph per batch)
        bt = make_batch(msa, ctx, meshes, models, counts, msa.MSA_FP64, args.tile, table, xi, args)
        kw = dict(run_kw)
        kw["tol"] = args.fp64_tol
        bt.run(**kw)
        ctx.synchronize()
        errs = []
        for c in range(len(models)):
            st = bt.stats(c)
            assert (st["status"] == msa.MSA_OK).all(), "tight FP64 solve did not converge"
            for s in range(counts[c]):
                ur = bt.solution(c, s)
                errs.append(np.linalg.norm(u_mixed[c][s] - ur) / np.linalg.norm(ur))
        max_err = float(np.max(errs))
        extra["rel_err_vs_fp64_gpu"] = dict(max=max_err, median=float(np.median(errs)),
                                            fp64_tol=args.fp64_tol)
        log(f"[bench] mixed vs tight FP64 batch: max rel err {max_err:.3e}")
        bt.close()

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
    for m in meshes:
        m.close()

    # ------------------------------------------------ the other configurations
    if not args.no_c3:
        extra["c3_fp64_batch_256x32"] = leg_c3(msa, ctx, args)
    if not args.no_c2:
        extra["c2_fp64_single_64"] = leg_c2(msa, ctx, args)
    if not args.no_c5:
        extra["c5_newmark"] = leg_newmark(msa, ctx, args, args.newmark_steps)
    ctx.close()

    # ------------------------------------------------ CPU baseline + oracle check
    cpu, indep = None, None
    if not args.no_cpu:
        firsts = [u_mixed[c][0] for c in range(len(sizes))]
        cpu, par, indep = cpu_baseline(args, sizes, counts, xi, firsts)
        extra["cpu_baseline_parallel"] = par
        extra["rel_err_vs_cpu_oracle"] = dict(per_size=indep, max=max(indep.values()),
                                              oracle_tol=args.oracle_tol,
                                              samples="first sample of each class")
        log(f"[bench] mixed vs CPU oracle: {indep}")

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
            "unroll": args.unroll,
            "spmv_form": ("matrix-free FP32 product from the cell moduli (differences from "
                          "the node's value); the stored upper triangle is still assembled"
                          if mf_all else
                          "symmetric: stored upper triangle, lower blocks as transposes "
                          "(the slot unroll does not apply)" if sym_all else "full storage"),
            "schedule_from": "autotuner" if tuned else "command line",
            "autotune_ms_per_sample": ([dict(tile=k[0], unroll=k[1], ms=v)
                                        for k, v in tuned["table"].items()] if tuned else None),
            "precond": args.precond, "matrix_cores": bool(args.matrix_cores),
            "graph_iters": args.graph_iters, "all_converged": bool(all_ok),
            "max_rel_err_vs_cpu_oracle": max(indep.values()) if indep else None,
            "max_rel_err_vs_fp64_gpu": max_err,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
        "iters_mean_per_size": {str(n): float(np.mean(iters_mixed[c])) for c, n in enumerate(sizes)},
        "outer_mean_per_size": {str(n): float(np.mean(outer_mixed[c])) for c, n in enumerate(sizes)},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
