"""SIMP solves/s on one MI355X, with the K.u roofline and a CPU baseline.

Workload (default): one 64x64x64 SIMP problem, hex8 elements, E0=1,
Emin=1e-9, nu=0.3, p=3, volfrac 0.3, rmin 1.5, clamp at x=0, 4 seeded unit
point loads on face x=max (seed 0), FP32 Jacobi-PCG (FP64 dots) at relative
tolerance 1e-6 with warm starts, 100 OC iterations plus the final compliance
solve. One step = one complete SIMP solve from the uniform initial density.

Roofline: the dominant kernel is k_cg_modal, one launch per Jacobi-PCG
iteration (K.z, the vector updates and the dot partials). Its algorithmic
bytes per launch are those of a minimal fused FP32 PCG iteration,
120*n_nodes + 4*n_elems: the K.u operand and result (24 B/node), the element
moduli E_e (4 B/element), and x, p, r, Ap, M^-1 read plus x, r, p written
(96 B/node). Its average duration comes from HIP events that bracket sampled
chunks of back-to-back launches on the solver's stream during the timed
steps (chunks that hold early-exit launches after convergence are dropped).

CPU baseline: the oracle's multi-core FP64 C restatement (oracle/simp_cpu.c,
matrix-free Jacobi-PCG, OpenMP over all host cores) on a bounded sample of
the same problem: the first OC iterations, timed; the per-PCG-iteration cost
is scaled by the PCG iteration count of the GPU's full solve.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "SIMP solves/sec & K·u GB/s vs HBM peak; batched p99 latency @1e-6 compliance"
SCHED_WARM_NI = 8
RECYCLE_BATCH = 6  # stored solutions per recycling job in the batch legs
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, nargs=3, default=[64, 64, 64])
    ap.add_argument("--oc-iters", type=int, default=100)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--recycle", type=int, default=6,
                    help="warm starts from the Galerkin combination of the last m solutions "
                         "(0: from the last solution alone)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true",
                    help="default workload: skip the side legs (accuracy against FP64, "
                         "reference-literal solve, 48-problem batch, serving)")
    ap.add_argument("--cpu-oc-iters", type=int, default=2)
    ap.add_argument("--cpu-scipy", action="store_true",
                    help="also time SciPy CSR Jacobi-PCG at the headline grid (about a minute "
                         "of assembly and iterations)")
    ap.add_argument("--traffic-file",
                    default=str(ROOT / "profiles" / "k_cg_modal_64x64x64_traffic.json"))
    ap.add_argument("--calib-mib", type=int, default=0,
                    help="first run one elementwise add over a float32 buffer of this many MiB "
                         "(a known-size kernel for scaling HBM counters)")
    ap.add_argument("--workload",
                    choices=["single64", "batch48", "serve", "mixed128", "ku", "evaluator",
                             "datagen"],
                    default="single64",
                    help="single64: one 64^3 problem (default); batch48: 48 mixed-size problems "
                         "on the device scheduler against the same problems one by one; serve: "
                         "open-loop request serving at 50/200/800 req/s; mixed128: 16-bit PCG "
                         "storage under FP64 refinement on 128x64x64, errors against FP64; ku: "
                         "K.u and one PCG iteration on a grid larger than the Infinity Cache, "
                         "beside a device copy; evaluator: a BO batch of latent vectors through "
                         "VoxelFEMEvaluator.evaluate_batch against evaluate() one by one; "
                         "datagen: TopoDataGenerator samples (SIMP + STL + JSON) as one "
                         "device batch against generate_single one by one")
    ap.add_argument("--ku-grid", type=int, nargs=3, default=[256, 256, 256])
    ap.add_argument("--mixed-oc-iters", type=int, default=50)
    ap.add_argument("--batch-n", type=int, default=48)
    ap.add_argument("--batch-oc-iters", type=int, default=60)
    ap.add_argument("--batch-legs", default="fp32,fp32_recycled,fp32+ir,fp32+ir_1e-9",
                    help="scheduler legs of the batch workload, comma-separated")
    ap.add_argument("--batch-only", action="store_true",
                    help="batch workload: the scheduler legs alone (no one-by-one legs, no "
                         "FP64 sample), e.g. for counter passes")
    ap.add_argument("--serve-n", type=int, default=60, help="requests per saturating rate")
    ap.add_argument("--serve-burst", type=int, default=24,
                    help="requests of the burst that measures the service capacity")
    ap.add_argument("--serve-seconds", type=float, default=15.0,
                    help="seconds of arrivals in each below-capacity leg")
    ap.add_argument("--serve-warm-compliance", action="store_true",
                    help="SIMP requests warm-start their compliance solves")
    ap.add_argument("--serve-recycle", type=int, default=0,
                    help="SIMP requests start warm OC solves from the Galerkin combination of "
                         "their last this many solutions (0: the last u)")
    ap.add_argument("--serve-ni", type=int, default=8, help="sweep depth of the serving slots")
    ap.add_argument("--serve-slots", type=int, nargs=2, default=[8, 4],
                    help="serving slots for 32^3 and for 64^3 requests")
    ap.add_argument("--serve-policy", choices=["fifo", "sjf"], default="sjf",
                    help="admission order: arrival (fifo) or single solves first (sjf)")
    ap.add_argument("--eval-q", type=int, default=32, help="latent vectors per BO batch")
    ap.add_argument("--eval-reps", type=int, default=5, help="timed BO batches")
    ap.add_argument("--datagen-n", type=int, default=64, help="generated samples per batch")
    return ap.parse_args()


def _line(metric, value, unit, steps, warmup, ms, dtype, config, **extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
           "warmup": warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}
    out.update(extra)
    return out


def batch48_legs(a, refine_leg=True, fp64_sample=4, refined_one_by_one=False):
    """C4: 48 SIMP problems, grids drawn by seed from {32^3, 48x24x24, 64x32x32, 64^3},
    volfrac ~ U[0.2, 0.4], 1-4 seeded point loads on face x = max, clamp at x = 0,
    60 OC iterations, all in one device scheduler (independent progress, ticks
    replayed from a HIP graph), timed beside the same problems run one after
    another through the single-problem FP32 path (PCG at 1e-6). Legs: the
    FP32 batch at 1e-6 with plain warm starts and with recycled starts (the
    Galerkin combination of each job's last 6 solutions), and the
    mixed-precision batch with recycled starts (FP64 refinement to 1e-7
    around FP32 inner solves at 1e-2, and to 1e-9 around inner solves at
    1e-3). One by one: plain FP32 (the single-problem path as the reference
    runs it) and FP32 with recycled starts (the same algorithm as the
    recycled batch). A sample of one problem per size class is also solved in
    FP64 at 1e-10: each leg's compliance and density error against it."""
    import torch
    from genpipeline_amd import Problem
    from genpipeline_amd.problems import batch_specs
    from genpipeline_amd.sched import run_batch

    ws = batch_specs(seed=0, n=a.batch_n, n_iters=a.batch_oc_iters)
    dev = torch.device("cuda", 0)
    Ps = [Problem(w.spec, w.fixed, w.f, device=dev) for w in ws]
    # workgroup shape and per-grid sweep depths from the project's autotuner
    # (cached for this mix of grid sizes; tuned on the spot when missing)
    from genpipeline_amd import autotune
    tile, nis = autotune.sched_shape(Ps)
    shape = {"tile": list(tile), "ni": {f"{P.spec.nx}x{P.spec.ny}x{P.spec.nz}": n
                                        for P, n in zip(Ps, nis)}}
    sched = dict(tile="auto", graph_ticks=32, max_iter=20000)
    legs = {"fp32": dict(tol=1e-6), "fp32_recycled": dict(tol=1e-6, recycle=RECYCLE_BATCH)}
    if refine_leg:
        legs["fp32+ir"] = dict(tol=1e-7, refine=True, inner_tol=1e-2, recycle=RECYCLE_BATCH)
        # tight enough for the 1e-6 compliance bar against FP64 at 1e-10
        legs["fp32+ir_1e-9"] = dict(tol=1e-9, refine=True, inner_tol=1e-3,
                                    recycle=RECYCLE_BATCH)
    want = [x for x in a.batch_legs.split(",") if x]
    unknown = [x for x in want if x not in legs]
    if unknown:
        raise SystemExit(f"unknown batch legs {unknown}; choose from {list(legs)}")
    legs = {k: v for k, v in legs.items() if k in want}
    run_batch(Ps[:4], 2, tile=tile, ni=SCHED_WARM_NI, graph_ticks=32, max_iter=20000)  # warm-up
    out = {}
    results = {}
    for name, kw in legs.items():
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res, ticks = run_batch(Ps, a.batch_oc_iters, **sched, **kw)
        torch.cuda.synchronize()
        t_b = time.perf_counter() - t0
        results[name] = res
        pcg_bytes = sum(r.pcg_iters * (120 * P.spec.n_nodes + 4 * P.spec.n_elem)
                        for r, P in zip(res, Ps))
        out[name] = {"seconds": t_b, "solves_per_s": len(Ps) / t_b, "ticks": ticks,
                     "recycle": kw.get("recycle", 0),
                     "us_per_tick": t_b / ticks * 1e6,
                     "pcg_iters_total": sum(r.pcg_iters for r in res),
                     "pcg_algorithmic_gbs": pcg_bytes / t_b / 1e9,
                     "worst_status": max(r.status for r in res)}
        print(f"[bench] batch48 scheduler {name}: {t_b:.2f} s, {ticks} ticks", file=sys.stderr,
              flush=True)
    if a.batch_only:
        for P in Ps:
            P.close()
        return {"problems": len(Ps), "oc_iters": a.batch_oc_iters, "scheduler_shape": shape,
                "legs": out, "scheduler_legs_only": True}
    t0 = time.perf_counter()
    seq = [P.run(a.batch_oc_iters, tol=1e-6, max_iter=20000, precision="fp32", warm_start=True)
           for P in Ps]
    torch.cuda.synchronize()
    t_s = time.perf_counter() - t0
    print(f"[bench] batch48 one by one (fp32, 1e-6): {t_s:.2f} s", file=sys.stderr, flush=True)
    one = {"seconds": t_s, "solves_per_s": len(Ps) / t_s,
           "pcg_iters_total": sum(r.pcg_iters for r in seq)}
    t0 = time.perf_counter()
    seq_rc = [P.run(a.batch_oc_iters, tol=1e-6, max_iter=20000, precision="fp32",
                    warm_start=True, recycle=RECYCLE_BATCH) for P in Ps]
    torch.cuda.synchronize()
    t_rc = time.perf_counter() - t0
    print(f"[bench] batch48 one by one (fp32, 1e-6, recycled): {t_rc:.2f} s", file=sys.stderr,
          flush=True)
    one_rc = {"seconds": t_rc, "solves_per_s": len(Ps) / t_rc,
              "pcg_iters_total": sum(r.pcg_iters for r in seq_rc)}
    if "fp32_recycled" in out:
        out["fp32_recycled"]["speedup_vs_fp32_recycled_one_by_one"] = (
            t_rc / out["fp32_recycled"]["seconds"])
    one_ir = None
    if refined_one_by_one:
        # the same accuracy one problem at a time: FP64 refinement to 1e-9
        # around FP32 inner solves, recycled starts (the fastest single path)
        t0 = time.perf_counter()
        seq_ir = [P.run(a.batch_oc_iters, tol=1e-9, max_iter=20000, precision="fp32",
                        refine=True, inner_tol=1e-3, warm_start=True, recycle=6) for P in Ps]
        torch.cuda.synchronize()
        t_ir = time.perf_counter() - t0
        print(f"[bench] batch48 one by one (fp32+ir 1e-9, recycled): {t_ir:.2f} s",
              file=sys.stderr, flush=True)
        one_ir = {"seconds": t_ir, "solves_per_s": len(Ps) / t_ir,
                  "pcg_iters_total": sum(r.pcg_iters for r in seq_ir)}
        if "fp32+ir_1e-9" in out:
            out["fp32+ir_1e-9"]["speedup_vs_fp32_ir_1e-9_one_by_one"] = (
                t_ir / out["fp32+ir_1e-9"]["seconds"])
    def _rel(name, ref):
        if name not in results:
            return None
        return max(abs(x.compliance - y.compliance) / abs(y.compliance)
                   for x, y in zip(results[name], ref))
    rel = _rel("fp32", seq)
    rel_rc = _rel("fp32_recycled", seq_rc)
    for name in out:
        out[name]["speedup_vs_fp32_one_by_one"] = t_s / out[name]["seconds"]
    # FP64 references for one problem of each size class
    sample, seen = [], set()
    for i, w in enumerate(ws):
        g = (w.spec.nx, w.spec.ny, w.spec.nz)
        if g not in seen and len(sample) < fp64_sample:
            seen.add(g)
            sample.append(i)
    err = {name: {"compliance_rel_max": 0.0, "density_max_abs": 0.0} for name in out}
    err["fp32_one_by_one"] = {"compliance_rel_max": 0.0, "density_max_abs": 0.0}
    for i in sample:
        ref = Ps[i].run(a.batch_oc_iters, tol=1e-10, max_iter=40000, precision="fp64",
                        warm_start=True)
        for name, rs in list(results.items()) + [("fp32_one_by_one", seq)]:
            e = err[name]
            e["compliance_rel_max"] = max(e["compliance_rel_max"],
                                          abs(rs[i].compliance - ref.compliance)
                                          / abs(ref.compliance))
            e["density_max_abs"] = max(e["density_max_abs"],
                                       float((rs[i].xPhys - ref.xPhys).abs().max()))
    for P in Ps:
        P.close()
    return {"problems": len(Ps), "oc_iters": a.batch_oc_iters, "scheduler_shape": shape, "legs": out,
            "fp32_one_by_one": one, "fp32_recycled_one_by_one": one_rc,
            "fp32_ir_1e-9_one_by_one": one_ir,
            "max_rel_compliance_diff_fp32_batch_vs_one_by_one": rel,
            "max_rel_compliance_diff_fp32_recycled_batch_vs_one_by_one": rel_rc,
            "error_vs_fp64_1e-10": err,
            "fp64_sample": [f"{ws[i].spec.nx}x{ws[i].spec.ny}x{ws[i].spec.nz}#{i}" for i in sample]}


def bench_batch48(a):
    """C4 as its own bench line (value: the FP32 batch with recycled starts,
    solves/s)."""
    d = batch48_legs(a, refined_one_by_one=True)
    leg = d["legs"].get("fp32_recycled") or next(iter(d["legs"].values()))
    return _line("SIMP solves/sec (48-problem batch, device scheduler)", leg["solves_per_s"],
                 "solves/s", 1, 1, leg["seconds"] * 1e3, "f32",
                 {"workload": "simp_batch48_c4_seed0", "problems": d["problems"],
                  "oc_iters": a.batch_oc_iters,
                  "grids": "32^3 / 48x24x24 / 64x32x32 / 64^3 by seed",
                  "pcg": "jacobi, fp32 vectors, fp64 dots, rel tol 1e-6, warm start from the "
                         f"Galerkin combination of each job's last {RECYCLE_BATCH} solutions; "
                         "fp32 leg: plain warm starts; fp32+ir legs: FP64 refinement to 1e-7 / "
                         "1e-9 around FP32 inner solves at 1e-2 / 1e-3, recycled starts",
                  "scheduler": "autotuned 256-lane workgroup shape and per-grid sweep depth "
                               "(extra.scheduler_shape), 32-tick HIP graph"},
                 extra=d)


def serve_legs(a, saturating=(50,)):
    """C5: Poisson arrivals of 32^3 (60%) / 64^3 (40%) density fields
    (smoothed noise, solid fraction U[0.15, 0.45]); half single solves
    (GPUVoxelFEM semantics, FP64-refined to 1e-8 around FP32 inner solves),
    half 30-iteration SIMP refinements (SIMPSolverGPU.run with x_init);
    continuous batching on the device scheduler. First the service capacity
    (a burst: every request arriving at t = 0), then open-loop arrivals at
    0.5 and 0.8 of that capacity (a stable queue: latency is service time plus
    queueing) with about a_serve_seconds of arrivals each, then the saturating
    rates named in the configuration (above capacity the latencies are drain
    times of the queue)."""
    import numpy as np
    import torch
    from genpipeline_amd.problems import serving_requests
    from genpipeline_amd.serve import Server, ServeConfig, latency_summary

    n32, n64 = a.serve_slots
    cfg = ServeConfig(policy=a.serve_policy, simp_warm_compliance=a.serve_warm_compliance,
                      simp_recycle=a.serve_recycle)
    srv = Server({(32, 32, 32): n32, (64, 64, 64): n64}, cfg=cfg, tile=(14, 18), ni=a.serve_ni,
                 chunk_ticks=64)
    srv.run(serving_requests(1000, 6, seed=99))  # warm-up

    def leg(lam, n, seed):
        reqs = serving_requests(lam, n, seed=seed)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = srv.run(reqs)
        wall = time.perf_counter() - t0
        out = latency_summary(res, wall)
        out["rate_per_s"] = lam
        for k in ("solve", "simp"):
            lat = [r.latency for r in res if r.kind == k]
            if lat:
                out[f"{k}_n"] = len(lat)
                out[f"{k}_worst_status"] = max(r.status for r in res if r.kind == k)
                out[f"{k}_p50_ms"] = float(np.percentile(lat, 50) * 1e3)
                out[f"{k}_p99_ms"] = float(np.percentile(lat, 99) * 1e3)
                out[f"{k}_pcg_iters_mean"] = float(np.mean([r.pcg_iters for r in res
                                                            if r.kind == k]))
        print(f"[bench] serve lam={lam:.3g} n={n}: {out}", file=sys.stderr, flush=True)
        return out, wall, len(res)

    per = {}
    burst, wall, n = leg(1e9, a.serve_burst, 1)
    per["burst"] = burst
    mu = burst["solves_per_s"]
    total_n, total_t = n, wall
    for frac in (0.5, 0.8):
        lam = frac * mu
        m = max(12, int(round(lam * a.serve_seconds)))
        per[f"{frac:.1f}x_capacity"], wall, n = leg(lam, m, int(frac * 10))
        total_n += n
        total_t += wall
    for lam in saturating:
        per[f"{lam}_per_s"], wall, n = leg(float(lam), a.serve_n, lam)
        total_n += n
        total_t += wall
    srv.close()
    return {"service_capacity_solves_per_s": mu, "per_rate": per,
            "solves_per_s_overall": total_n / total_t, "slots": {"32^3": n32, "64^3": n64},
            "admission": a.serve_policy, "simp_warm_compliance": a.serve_warm_compliance,
            "simp_recycle": a.serve_recycle, "peak_problem_bytes": srv.peak_used}


def bench_serve(a):
    """C5 as its own bench line (value: solves/s over every leg), with the
    configuration's three arrival rates 50 / 200 / 800 per second."""
    d = serve_legs(a, saturating=(50, 200, 800))
    return _line("SIMP/FEM request serving: solves/sec with p50/p99 latency",
                 d["solves_per_s_overall"], "solves/s", len(d["per_rate"]), 1, 0.0, "f32",
                 {"workload": "serve_c5_32cube_64cube", "requests_per_saturating_rate": a.serve_n,
                  "burst_requests": a.serve_burst, "stable_leg_seconds": a.serve_seconds,
                  "slots": d["slots"], "scheduler": f"14x18 tile, ni {a.serve_ni}",
                  "admission": a.serve_policy},
                 extra=d)


def bench_evaluator(a):
    """BO evaluator batch: q latent vectors (N(0,1), 32-dim) decoded to 64^3
    fields by a seeded stand-in decoder (no VAE checkpoint exists here),
    down-sampled and smoothed to 32^3 on the device, and solved with
    GPUVoxelFEM semantics (x_min clamp, 1000 N over the x_max face in -z) as
    one scheduler batch; timed beside evaluate() per vector (device
    pre-processing, FP64-refined GPUVoxelFEM solve), and the device
    pre-processing beside the reference's SciPy chain on the host."""
    import torch
    from genpipeline_amd.fem import VoxelFEMEvaluator
    from genpipeline_amd.fem.voxel_fem import decode_batch
    from genpipeline_amd.fem.voxel_prep import prep_voxels
    from genpipeline_amd.pipeline_utils import smooth_voxels
    from genpipeline_amd.problems import SyntheticDecoder, bo_latents
    from scipy.ndimage import zoom

    dev = torch.device("cuda", 0)
    dec = SyntheticDecoder(seed=0, device=dev)
    q = a.eval_q
    ev = VoxelFEMEvaluator(vae_model=dec, use_gpu=True)
    ev.batch_slots = min(q, 32)
    ev.evaluate_latents(bo_latents(q, seed=99))  # warm-up: scheduler, graphs, problems

    def timed(refine):
        ev.batch_refine = refine
        times, outs = [], []
        for rep in range(a.eval_reps):
            zs = bo_latents(q, seed=rep)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            outs.append(ev.evaluate_latents(zs))
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            print(f"[bench] evaluator batch {rep} (refine {refine}): {times[-1] * 1e3:.1f} ms",
                  file=sys.stderr, flush=True)
        return times, outs

    times_u, outs_u = timed(False)
    times, outs = timed(True)
    t_b = sum(times)
    # one by one: the reference's per-vector evaluate() on the same first batch
    single = VoxelFEMEvaluator(vae_model=dec, use_gpu=True)
    zs = bo_latents(q, seed=0)
    single.evaluate(zs[0], dec)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ref = [single.evaluate(z, dec) for z in zs]
    torch.cuda.synchronize()
    t_s = time.perf_counter() - t0
    rel = max(abs(o["compliance"] - r["compliance"]) / r["compliance"]
              for o, r in zip(outs[0], ref))
    rel_u = max(abs(o["compliance"] - r["compliance"]) / r["compliance"]
                for o, r in zip(outs_u[0], ref))
    # pre-processing alone: device kernels for the batch vs SciPy per field
    raw = decode_batch(list(zs), dec)
    prep_voxels(raw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        prep_voxels(raw)
    torch.cuda.synchronize()
    t_prep = (time.perf_counter() - t0) / 10
    host = raw.cpu().numpy()
    n_cpu = min(q, 8)
    t0 = time.perf_counter()
    for b in range(n_cpu):
        smooth_voxels(zoom(host[b], 0.5, order=1))
    t_prep_cpu = (time.perf_counter() - t0) / n_cpu * q
    ev.close()
    single.close()
    n = q * a.eval_reps
    return _line("FEM evaluations/sec (BO batch through VoxelFEMEvaluator.evaluate_batch)",
                 n / t_b, "evaluations/s", a.eval_reps, 1, t_b / a.eval_reps * 1e3, "f32",
                 {"workload": f"bo_evaluator_q{q}_64cube_decode_32cube_fem",
                  "latent_dim": 32, "decoded": [64, 64, 64], "fem_grid": [32, 32, 32],
                  "solve": "GPUVoxelFEM semantics and refinement: FP32 Jacobi-PCG (FP64 dots) "
                           "batches at 1e-3, FP64 residual passes to 1e-8; "
                           f"{ev.batch_slots} scheduler slots, 14x18 tile, ni 8"},
                 extra={"one_by_one_evaluations_per_s": q / t_s,
                        "speedup_vs_one_by_one": (n / t_b) / (q / t_s),
                        "max_rel_compliance_vs_one_by_one": rel,
                        "unrefined_evaluations_per_s": n / sum(times_u),
                        "unrefined_max_rel_compliance_vs_one_by_one": rel_u,
                        "prep_device_ms_per_batch": t_prep * 1e3,
                        "prep_scipy_host_ms_per_batch": t_prep_cpu * 1e3,
                        "batch_ms": [t * 1e3 for t in times]})


def bench_datagen(a):
    """Data generation: TopoDataGenerator samples with the reference's defaults
    (medium grids, organic geometry mix, 30 OC iterations, penal 3.5,
    DenseGPUSolver semantics), drawn from numpy seed 0: SIMP, density_to_stl
    and the JSON record, as one device batch (generate) against the same
    number of samples through generate_single one after another."""
    import tempfile

    import numpy as np
    import torch
    from genpipeline_amd.topology.topo_data_gen import TopoDataGenerator

    n = a.datagen_n
    with tempfile.TemporaryDirectory() as tmp:
        g = TopoDataGenerator(output_dir=tmp)
        np.random.seed(99)
        g.generate(4)  # warm-up
        np.random.seed(0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        stems = g.generate(n)
        torch.cuda.synchronize()
        t_b = time.perf_counter() - t0
        n_one = min(n, 16)
        np.random.seed(0)
        t0 = time.perf_counter()
        for i in range(n_one):
            try:
                g.generate_single(i, n_one)
            except ValueError:
                pass
        torch.cuda.synchronize()
        t_s = (time.perf_counter() - t0) / n_one * n
    print(f"[bench] datagen batch {t_b:.2f} s for {n}; one by one {t_s:.2f} s (from {n_one})",
          file=sys.stderr, flush=True)
    return _line("SIMP dataset samples/sec (TopoDataGenerator.generate, one device batch)",
                 n / t_b, "samples/s", 1, 1, t_b * 1e3, "f32",
                 {"workload": f"topo_datagen_{n}_medium_organic_seed0", "samples": n,
                  "oc_iters": 30, "semantics": "DenseGPUSolver (reference default use_gpu=True)",
                  "outputs": "density_to_stl STL + JSON per sample"},
                 extra={"one_by_one_samples_per_s": n / t_s, "speedup_vs_one_by_one": t_s / t_b,
                        "written": len(stems), "one_by_one_timed_on": n_one})


def bench_mixed128(a):
    """C3: 128x64x64 (524,288 elements), distributed traction of total magnitude
    1 on face x = max, density from seeded smoothed noise (sigma 2) clipped to
    [0.01, 1] and rescaled to volfrac 0.25; SIMP from that density with FP64
    refinement to an outer residual of 1e-6 around an inner PCG at 1e-3 whose
    iterate and search direction are stored in 16 bits (fp16, the value of
    the line, and bf16), next to the FP32-inner refinement, plain FP32 PCG and
    the FP64 reference run. The 16-bit and FP32 runs start each solve from
    the Galerkin combination of the last --recycle solutions (0: the last one
    alone). The compliance and density errors of each are reported against
    FP64."""
    per, w = mixed128_legs(a)
    dt = per["fp16+ir"]["seconds"]
    return _line("SIMP solves/sec, 16-bit PCG storage + FP64 refinement (128x64x64)", 1.0 / dt,
                 "solves/s", 1, 1, dt * 1e3, "fp16-storage/f32-arith/f64-refine",
                 {"workload": f"simp_mixed_128x64x64_seed{a.seed}", "grid": [128, 64, 64],
                  "oc_iters": w.n_iters, "outer_tol": 1e-6, "inner_tol": 1e-3,
                  "recycle": a.recycle,
                  "load": "traction, total 1, face x=max", "volfrac": 0.25},
                 extra={"runs": per})


def mixed128_legs(a, names=None):
    """The C3 runs named in names (all five when None; the FP64 reference
    always runs first): per-run seconds, PCG iterations and errors against
    FP64, and the workload."""
    import torch
    from genpipeline_amd import Problem
    from genpipeline_amd.problems import mixed_128x64x64

    w = mixed_128x64x64(seed=a.seed, n_iters=a.mixed_oc_iters)
    dev = torch.device("cuda", 0)
    P = Problem(w.spec, w.fixed, w.f, device=dev)
    base = dict(x0=w.x0, xphys0=w.x0, max_iter=20000, max_outer=60, warm_start=True)
    rc = dict(recycle=a.recycle)
    runs = [("fp64", dict(precision="fp64", tol=1e-10)),
            ("fp16+ir", dict(precision="fp16", refine=True, tol=1e-6, inner_tol=1e-3, **rc)),
            ("bf16+ir", dict(precision="bf16", refine=True, tol=1e-6, inner_tol=1e-3, **rc)),
            ("fp32+ir", dict(precision="fp32", refine=True, tol=1e-6, inner_tol=1e-3, **rc)),
            ("fp32", dict(precision="fp32", tol=1e-6, **rc))]
    if names is not None:
        runs = [r for r in runs if r[0] == "fp64" or r[0] in names]
    # warm-up: the 16-bit path when it runs
    P.run(2, **base, **dict(runs).get("fp16+ir", runs[0][1]))
    res = {}
    for name, kw in runs:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = P.run(w.n_iters, **base, **kw)
        torch.cuda.synchronize()
        res[name] = (r, time.perf_counter() - t0)
        print(f"[bench] mixed128 {name}: {res[name][1]:.2f} s, {r.pcg_iters} PCG its, "
              f"C={r.compliance:.12g}", file=sys.stderr, flush=True)
    ref = res["fp64"][0]
    per = {}
    for name, (r, dt) in res.items():
        per[name] = {"seconds": dt, "solves_per_s": 1.0 / dt, "pcg_iters": r.pcg_iters,
                     "status": r.status, "compliance": r.compliance,
                     "compliance_rel_err_vs_fp64": abs(r.compliance - ref.compliance)
                     / abs(ref.compliance),
                     "density_max_abs_err_vs_fp64": float((r.xPhys - ref.xPhys).abs().max())}
    P.close()
    return per, w


def _event_us(fn, reps):
    """Average microseconds of fn() over reps back-to-back calls, from HIP
    events on the current stream (fn launches on that stream)."""
    import torch
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) * 1e3 / reps


def bench_ku(a):
    """K.u on a grid whose operands do not fit the 256 MB Infinity Cache (default
    256^3: u and y are 204 MB each in FP32, E 67 MB), so it streams from HBM.
    Reported beside one fused PCG iteration on the same grid and a 1 GiB device
    copy (the bandwidth HBM delivers to a plain read + write stream). Every
    modal tile and sweep depths 4/8/16 (K.u) and 8/12/16/24 (PCG iteration,
    HIP-event timed per launch) are tried; the best is reported."""
    import numpy as np
    import torch
    from genpipeline_amd import Problem, ProblemSpec, autotune
    from genpipeline_amd.bcs import clamp_x0

    dev = torch.device("cuda", 0)
    n = (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    copy_us = _event_us(lambda: dst.copy_(src), 20)
    copy_gbs = 2 * 4 * n / (copy_us * 1e-6) / 1e9
    del src, dst
    nx, ny, nz = a.ku_grid
    spec = ProblemSpec(nx, ny, nz)
    f = np.zeros(spec.n_dof)
    f[-1] = -1.0
    P = Problem(spec, clamp_x0(nx, ny, nz), f, device=dev)
    rng = np.random.default_rng(0)
    E = torch.from_numpy(rng.uniform(1e-3, 1.0, spec.n_elem).astype(np.float32)).to(dev)
    u = torch.from_numpy(rng.standard_normal(spec.n_dof).astype(np.float32)).to(dev)
    ku_bytes = 24 * spec.n_nodes + 4 * spec.n_elem
    it_bytes = 120 * spec.n_nodes + 4 * spec.n_elem
    f_dense = autotune.dense_rhs(P)  # dense PCG vectors from the first iteration
    ku_t, it_t = {}, {}
    for tj, tk in autotune.TILES:
        for ni in (4, 8, 16):
            P.set_ku_tile(tj, tk, ni, precision="fp32")
            ku_t[(tj, tk, ni)] = _event_us(lambda: P.ku(E, u), 10)
        for ni in (8, 12, 16, 24):
            it_t[(tj, tk, ni)] = autotune.time_pcg(P, E, (tj, tk, ni), "fp32", iters=120,
                                                   repeats=2, f=f_dense)
    kb = min(ku_t, key=ku_t.get)
    ib = min(it_t, key=it_t.get)
    P.close()
    ku_gbs = ku_bytes / (ku_t[kb] * 1e-6) / 1e9
    it_gbs = it_bytes / (it_t[ib] * 1e-6) / 1e9
    print(f"[bench] ku {nx}x{ny}x{nz}: copy {copy_gbs:.0f} GB/s; K.u {kb} {ku_t[kb]:.1f} us "
          f"{ku_gbs:.0f} GB/s; PCG iteration {ib} {it_t[ib]:.1f} us {it_gbs:.0f} GB/s",
          file=sys.stderr, flush=True)
    return _line("K·u GB/s vs HBM peak (HBM-resident grid)", ku_gbs, "GB/s", 10, 1,
                 ku_t[kb] * 1e-3, "f32",
                 {"workload": f"ku_{nx}x{ny}x{nz}_modal", "grid": [nx, ny, nz],
                  "ku_tile": list(kb), "pcg_tile": list(ib)},
                 roofline={"bound": "hbm", "kernel": "k_ku_modal<float> (y = K(E) u, fixed DOFs "
                           "zeroed)", "achieved": ku_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ku_gbs / HBM_PEAK_GBS, "traffic": None,
                           "algorithmic_bytes_per_launch": ku_bytes, "avg_launch_us": ku_t[kb]},
                 extra={"device_copy_gbs": copy_gbs, "device_copy_frac_of_peak":
                        copy_gbs / HBM_PEAK_GBS, "ku_frac_of_copy": ku_gbs / copy_gbs,
                        "pcg_iteration_us": it_t[ib], "pcg_iteration_gbs": it_gbs,
                        "pcg_iteration_frac_of_peak": it_gbs / HBM_PEAK_GBS,
                        "pcg_algorithmic_bytes": it_bytes,
                        "ku_us_by_tile": {f"{k[0]}x{k[1]}x{k[2]}": v for k, v in ku_t.items()},
                        "pcg_us_by_tile": {f"{k[0]}x{k[1]}x{k[2]}": v for k, v in it_t.items()}})


def cpu_baseline(w, gpu_pcg_iters, gpu_oc_iters, n_sample):
    import numpy as np
    from oracle import cpu_ref, simp_oracle as O

    threads = len(os.sched_getaffinity(0))
    KE = O.ke_hex8(w.spec.nu)
    grid = (w.spec.nx, w.spec.ny, w.spec.nz)
    t0 = time.perf_counter()
    x, xp, u, comp, its = cpu_ref.simp_run(grid, KE, w.f, w.fixed, n_sample, w.spec.volfrac, w.tol,
                                           law=w.spec.law, threads=threads)
    dt = time.perf_counter() - t0
    # the sample ran n_sample OC iterations and one final solve: n_sample+1 solves
    per_pcg_iter = dt / max(its, 1)
    est_solve_s = per_pcg_iter * gpu_pcg_iters
    return {
        "value": 1.0 / est_solve_s,
        "unit": "solves/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle/simp_cpu.c FP64 matrix-free Jacobi-PCG, OpenMP x{threads}, first "
                   f"{n_sample} OC iterations + final solve of the same {grid[0]}x{grid[1]}x{grid[2]} "
                   f"problem ({its} PCG iterations, {dt:.2f} s); per-PCG-iteration time scaled by the "
                   f"GPU solve's {gpu_pcg_iters} PCG iterations over {gpu_oc_iters} OC iterations"),
        "sample_seconds": dt,
        "sample_pcg_iters": its,
        "cpu_model": _cpu_model(),
        "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"),
    }


def accuracy_vs_fp64(P, w, r_head, plain_kw=None, recycle=0):
    """The headline solve's error against the same SIMP run in FP64 at PCG
    tolerance 1e-10 (compliance relative, density max-abs), and the error and
    rate of FP64 refinement around FP32 inner solves at an outer tolerance of
    1e-9, the mode that meets 1e-6 / 1e-5 (with the headline's recycled
    starts). plain_kw: also the FP32 run with plain warm starts (each solve
    from the last u alone)."""
    import torch
    out = {}
    ref = None
    legs = [("fp64_tol1e-10", dict(precision="fp64", tol=1e-10)),
            ("fp32+ir_tol1e-9", dict(precision="fp32", refine=True, tol=1e-9, inner_tol=1e-3,
                                     recycle=recycle))]
    if plain_kw is not None:
        legs.append(("fp32_plain_warm_start", plain_kw))
    for name, kw in legs:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rr = P.run(w.n_iters, **{**dict(max_iter=40000, warm_start=True), **kw})
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if ref is None:
            ref = rr
        out[name] = {"seconds": dt, "solves_per_s": 1.0 / dt, "pcg_iters": rr.pcg_iters,
                     "compliance": rr.compliance,
                     "compliance_rel_err_vs_fp64": abs(rr.compliance - ref.compliance)
                     / abs(ref.compliance),
                     "density_max_abs_err_vs_fp64": float((rr.xPhys - ref.xPhys).abs().max())}
        print(f"[bench] {name}: {dt:.2f} s, C={rr.compliance:.12g}", file=sys.stderr, flush=True)
    out["headline_compliance_rel_err_vs_fp64"] = (abs(r_head.compliance - ref.compliance)
                                                  / abs(ref.compliance))
    out["headline_density_max_abs_err_vs_fp64"] = float((r_head.xPhys - ref.xPhys).abs().max())
    return out


def cpu_reference_literal_c1():
    """The reference CPU solver's own method on its CPU-runnable configuration
    (32x16x16, the reference's element table, clamp law, SciPy SuperLU, one
    thread by construction; simp_solver.py:195-213), through the oracle:
    one direct solve at the uniform start density, timed. A run of C1 is 51
    such solves (50 OC iterations and the final compliance solve)."""
    from oracle import simp_oracle as O
    nx, ny, nz = 32, 16, 16
    P = O.reference_problem(nx, ny, nz, volfrac=0.3, ke="ref", law="clamp", oc="cpu")
    x = np.full(P.n_elem, 0.3)
    P.setup()
    t0 = time.perf_counter()
    u, _ = P.solve_u(x, "direct")
    dt = time.perf_counter() - t0
    return {"seconds_per_solve": dt, "solves_per_run": 51, "runs_per_s": 1.0 / (51 * dt),
            "cores": 1, "kind": "reference-literal (oracle: reference table, SciPy SuperLU)",
            "grid": [nx, ny, nz]}, u


REF_TABLE = Path(__file__).resolve().parent / "tests" / "golden" / "ref_simp_golden.npz"


def gpu_reference_literal_c1(n_solves=3):
    """Reference-literal mode on the device at the reference's CPU-runnable
    configuration (32x16x16, the reference's own element table -- the data
    of tests/golden -- clamp law, default cantilever supports and load):
    exact FP64 banded-LU solves (gp_direct) at the uniform start density,
    one warm-up solve and n_solves timed ones. A run of the configuration is
    51 such solves. Returns (figures, u)."""
    import torch
    from genpipeline_amd.topology import SIMPSolver
    from genpipeline_amd.topology._common import cached_problem
    with np.load(REF_TABLE, allow_pickle=False) as d:
        table = np.array(d["ke_ref_nu0.3"])
    nx, ny, nz = 32, 16, 16
    s = SIMPSolver(nx, ny, nz, ke=table)
    f, fixed = s._get_bcs(1.0)
    P = cached_problem(s, "clamp", "cpu", f, fixed, None, 0.3, ke=s._ke_dev)
    x = torch.full((nx * ny * nz,), 0.3, dtype=torch.float64, device=P.device)
    u, res = P.direct(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_solves):
        u, res = P.direct(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / n_solves
    out = {"seconds_per_solve": dt, "solves_per_run": 51, "runs_per_s": 1.0 / (51 * dt),
           "status": res.status, "grid": [nx, ny, nz],
           "kind": "reference-literal on the device (reference table, FP64 banded LU with "
                   "partial pivoting, one launch per free DOF)"}
    u = u.cpu().numpy()
    P.close()
    return out, u


def cpu_scipy_pcg(w, gpu_pcg_iters, n_iter=30):
    """The configuration-literal CPU method at the headline grid: FP64
    Jacobi-PCG on SciPy's assembled CSR K (the oracle's assembly, the
    reference's SpMV path; one core), timed over n_iter PCG iterations at
    the uniform start density and scaled by the GPU solve's PCG count.
    Assembly is reported separately as set-up."""
    import scipy.sparse.linalg as spla
    from oracle import simp_oracle as O
    sp = w.spec
    KE = O.ke_hex8(sp.nu)
    ed = O.edof_matrix(sp.nx, sp.ny, sp.nz)
    Ee = O.stiffness(np.full(sp.n_elem, sp.volfrac), sp.law, sp.penal, sp.E0, sp.Emin)
    t0 = time.perf_counter()
    K = O.assemble_K(KE, Ee, ed, sp.n_dof)
    free = np.ones(sp.n_dof, bool)
    free[w.fixed] = False
    Kf = K[free][:, free].tocsr()
    t_setup = time.perf_counter() - t0
    d = 1.0 / Kf.diagonal()
    b = w.f[free]
    M = spla.LinearOperator(Kf.shape, matvec=lambda r: d * r)
    t0 = time.perf_counter()
    spla.cg(Kf, b, rtol=1e-30, maxiter=n_iter, M=M)
    dt = time.perf_counter() - t0
    per = dt / n_iter
    return {"seconds_per_pcg_iter": per, "setup_seconds": t_setup, "sample_pcg_iters": n_iter,
            "solves_per_s": 1.0 / (per * gpu_pcg_iters), "cores": 1,
            "kind": "config-literal (SciPy CSR Jacobi-PCG, FP64)",
            "sample": f"{n_iter} scipy.sparse.linalg.cg iterations on the assembled free-DOF CSR "
                      f"K at uniform density, scaled by the GPU solve's {gpu_pcg_iters} PCG "
                      "iterations"}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def calibration_add(mib):
    """A known-size kernel for counter passes: one elementwise add that reads
    and writes mib MiB of float32."""
    import torch
    n = mib * (1 << 20) // 4
    ca = torch.ones(n, dtype=torch.float32, device=torch.device("cuda", 0))
    cb = torch.empty_like(ca)
    torch.add(ca, 1.0, out=cb)
    torch.cuda.synchronize()
    print(f"[bench] calibration add: read {4 * n} B, wrote {4 * n} B", file=sys.stderr,
          flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); independent SIMP "
              "problems shard one per GPU, and no multi-GPU figure is measured in this build.",
              file=sys.stderr)
        sys.exit(2)
    if a.calib_mib:
        calibration_add(a.calib_mib)
    if a.workload == "batch48":
        print(json.dumps(bench_batch48(a)), flush=True)
        return
    if a.workload == "serve":
        print(json.dumps(bench_serve(a)), flush=True)
        return
    if a.workload == "mixed128":
        print(json.dumps(bench_mixed128(a)), flush=True)
        return
    if a.workload == "ku":
        print(json.dumps(bench_ku(a)), flush=True)
        return
    if a.workload == "evaluator":
        print(json.dumps(bench_evaluator(a)), flush=True)
        return
    if a.workload == "datagen":
        print(json.dumps(bench_datagen(a)), flush=True)
        return
    import torch
    from genpipeline_amd import Problem
    from genpipeline_amd.problems import point_loads

    nx, ny, nz = a.grid
    w = point_loads(nx, ny, nz, seed=a.seed, n_loads=4, volfrac=0.3, n_iters=a.oc_iters, tol=1e-6)
    dev = torch.device("cuda", 0)
    P = Problem(w.spec, w.fixed, w.f, device=dev)
    run_kw = dict(tol=w.tol, max_iter=20000, precision="fp32", warm_start=True,
                  recycle=a.recycle)
    for k in range(a.warmup):
        rw = P.run(w.n_iters, **run_kw)
        print(f"[bench] warmup {k}: {rw.seconds:.2f} s, {rw.pcg_iters} PCG its, "
              f"C={rw.compliance:.6g}", file=sys.stderr, flush=True)
    ku_tile = P.ku_tile("fp32")
    print(f"[bench] K.u tile (tj, tk, ni, modal) = {ku_tile}", file=sys.stderr, flush=True)
    P.set_profiling(1)
    P.ku_profile(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results = []
    for k in range(a.steps):
        results.append(P.run(w.n_iters, **run_kw))
        print(f"[bench] step {k}: {results[-1].seconds:.2f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ku_ms, ku_samples = P.ku_profile(reset=True)
    P.set_profiling(0)
    r = results[-1]
    sp = w.spec
    ku_bytes = 24 * sp.n_nodes + 4 * sp.n_elem
    it_bytes = 120 * sp.n_nodes + 4 * sp.n_elem
    achieved = it_bytes / (ku_ms * 1e-3) / 1e9
    traffic = None
    tf = Path(a.traffic_file)
    if tf.exists():
        try:
            t = json.loads(tf.read_text())
            # the measured traffic holds for the grid and sweep tile it was taken on
            if t.get("grid") == list(a.grid) and t.get("tile") == list(ku_tile)[:3]:
                traffic = t.get("bytes_per_launch")
        except (ValueError, OSError):
            traffic = None
    out = {
        "metric": METRIC,
        "value": a.steps / dt,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"simp_{nx}x{ny}x{nz}_point_loads4_seed{a.seed}",
            "grid": [nx, ny, nz],
            "oc_iters": w.n_iters,
            "pcg": ("jacobi, fp32 vectors, fp64 dots, rel tol 1e-6, warm start"
                    + (f" from the Galerkin combination of the last {a.recycle} solutions"
                       if a.recycle else "")),
            "material": "hex8 E0=1 Emin=1e-9 nu=0.3 p=3 volfrac=0.3 rmin=1.5",
            "parallelism": "single problem, one GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_cg_modal<float> (one Jacobi-PCG iteration: K.z + vector updates + dots)",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": it_bytes,
            # K.z's own operand bytes (z read, w written, E read) inside the
            # fused launch, over the whole launch; K.u alone: --workload ku
            "ku_operand_bytes_per_launch": ku_bytes,
            "ku_operand_gbs_within_fused_launch": ku_bytes / (ku_ms * 1e-3) / 1e9,
            "avg_launch_us": ku_ms * 1e3,
            "event_samples": ku_samples,
        },
        "extra": {
            "compliance": r.compliance,
            "pcg_iters_per_solve": r.pcg_iters,
            "bisect_steps_per_solve": r.bisect_steps,
            "status": r.status,
            "solve_seconds_lib": [x.seconds for x in results],
            "ku_tile_fp32": list(ku_tile),
        },
    }
    if not a.no_extra:
        print("[bench] accuracy against FP64", file=sys.stderr, flush=True)
        plain_kw = dict(run_kw, recycle=0) if a.recycle else None
        out["extra"]["accuracy"] = accuracy_vs_fp64(P, w, r, plain_kw, a.recycle)
    P.close()
    u_lit = None
    if not a.no_extra:
        print("[bench] reference-literal mode on the device", file=sys.stderr, flush=True)
        try:
            out["extra"]["reference_literal_c1"], u_lit = gpu_reference_literal_c1()
        except Exception as e:
            out["extra"]["reference_literal_c1"] = {"error": f"{type(e).__name__}: {e}"}
        for key, fn in (("batch48", lambda: batch48_legs(a)),
                        ("mixed128", lambda: mixed128_legs(a, ("fp16+ir", "fp32+ir"))[0]),
                        ("serve", lambda: serve_legs(a))):
            try:
                out["extra"][key] = fn()
            except Exception as e:  # the headline line is printed whatever a side leg does
                out["extra"][key] = {"error": f"{type(e).__name__}: {e}"}
    if not a.no_cpu:
        print("[bench] cpu baseline sample", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, r.pcg_iters, w.n_iters, a.cpu_oc_iters)
        try:
            lit, u_ref = cpu_reference_literal_c1()
            out["cpu_baseline"]["reference_literal_c1"] = lit
            g = out["extra"].get("reference_literal_c1", {})
            if u_lit is not None and "seconds_per_solve" in g:
                # the same solve on both sides: SuperLU (CPU) against the device
                g["max_rel_diff_vs_superlu"] = float(np.abs(u_lit - u_ref).max()
                                                     / np.abs(u_ref).max())
                g["speedup_vs_superlu"] = lit["seconds_per_solve"] / g["seconds_per_solve"]
        except Exception as e:
            out["cpu_baseline"]["reference_literal_c1"] = {"error": f"{type(e).__name__}: {e}"}
        if a.cpu_scipy:
            out["cpu_baseline"]["scipy_csr_pcg"] = cpu_scipy_pcg(w, r.pcg_iters)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
