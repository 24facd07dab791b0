"""Autotuner for the modal K.u workgroup shapes.

Single problems: for a grid size and precision it times a fixed number of PCG
iterations (tolerance 0, so every candidate does the same work) for each
supported (tj, tk) tile and a range of sweep depths ni, and keeps the fastest.
Results are cached per (device ISA, grid, precision) in ``tuned_tiles.json``
next to this file; ``Problem`` consults the cache when it is created.

Device scheduler: for a mix of grid sizes it picks the 256-lane workgroup
shape shared by every slot and the sweep depth of each grid size by timing the
batch itself (``tune_sched``); ``sched.run_batch(tile="auto")`` uses it.

``python -m genpipeline_amd.autotune --grid 64 64 64 --batch48`` tunes on the
current GPU and rewrites the cache.
"""
from __future__ import annotations

import json
import threading
import time
from pathlib import Path

import numpy as np
import torch

CACHE = Path(__file__).resolve().parent / "tuned_tiles.json"
TILES = [(8, 32), (16, 16), (14, 18), (6, 34), (14, 34), (4, 64), (10, 26), (7, 34)]
NIS = [1, 2, 3, 4, 6, 8, 10, 11, 12, 16, 24, 32]
_lock = threading.Lock()
_mem = None


def _device_key(device=None) -> str:
    """The device's ISA name (e.g. ``gfx950``): stable across drivers and
    profilers, unlike the marketing name."""
    try:
        arch = torch.cuda.get_device_properties(device).gcnArchName
        return arch.split(":")[0] or "unknown"
    except Exception:
        return "unknown"


def _load() -> dict:
    global _mem
    if _mem is None:
        try:
            _mem = json.loads(CACHE.read_text()) if CACHE.exists() else {}
        except (OSError, ValueError):
            _mem = {}
    return _mem


def key(grid, precision, device=None) -> str:
    nx, ny, nz = grid
    return f"{_device_key(device)}|{nx}x{ny}x{nz}|{precision}"


def lookup(grid, precision="fp32", device=None):
    """Cached (tj, tk, ni) for this grid, or None."""
    with _lock:
        v = _load().get(key(grid, precision, device))
    return tuple(v["tile"]) if v else None


def store(grid, precision, tile, timings, device=None, persist=True):
    with _lock:
        d = _load()
        d[key(grid, precision, device)] = {"tile": list(tile), "us_per_iter": timings[tuple(tile)],
                                           "candidates": len(timings)}
        if persist:
            CACHE.write_text(json.dumps(d, indent=1, sort_keys=True))


def candidates(nx):
    out = []
    for tj, tk in TILES:
        for ni in NIS:
            if ni <= nx + 1:
                out.append((tj, tk, ni))
    return out


def dense_rhs(P, seed=0) -> torch.Tensor:
    """A seeded standard-normal load on every DOF (float64, on P's device):
    the PCG vectors are dense from the first iteration, as in the warm-started
    solves of a SIMP run. (From a point load, early iterations leave most
    vector entries exactly zero, which runs faster than dense data.)"""
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn(P.spec.n_dof, dtype=torch.float64, generator=g).to(P.device)


def time_pcg(P, E, tile, precision="fp32", iters=60, repeats=2, f=None) -> float:
    """Microseconds per PCG iteration launch with the given tile (tol = 0,
    load f, default dense_rhs(P)): HIP events around chunks of back-to-back
    launches on the solver's stream (the set-up and the status read-backs
    left out); the wall clock per iteration where no chunk was timed."""
    if f is None:
        f = dense_rhs(P)
    P.set_ku_tile(*tile, precision=precision)
    P.pcg(E, f=f, tol=0.0, max_iter=4, precision=precision)  # warm the code path
    best = float("inf")
    P.set_profiling(1)
    try:
        for _ in range(repeats):
            P.ku_profile(reset=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, r = P.pcg(E, f=f, tol=0.0, max_iter=iters, precision=precision)
            torch.cuda.synchronize()
            ms, n = P.ku_profile(reset=True)
            dt = ms * 1e-3 if n > 0 else (time.perf_counter() - t0) / max(r.iters, 1)
            best = min(best, dt)
    finally:
        P.set_profiling(0)
    return best * 1e6


def tune(P, precision="fp32", cands=None, iters=60, repeats=2, persist=True, log=None):
    """Times every candidate on problem P and stores the fastest tile."""
    sp = P.spec
    rng = np.random.default_rng(0)
    E = torch.from_numpy(rng.uniform(1e-3, 1.0, sp.n_elem).astype(np.float32)).to(P.device)
    cands = cands or candidates(sp.nx)
    f = dense_rhs(P)
    timings = {}
    for c in cands:
        timings[tuple(c)] = time_pcg(P, E, c, precision, iters, repeats, f)
        if log:
            log(f"[autotune] {sp.nx}x{sp.ny}x{sp.nz} {precision} tile={c} "
                f"{timings[tuple(c)]:.2f} us/iter")
    best = min(timings, key=timings.get)
    P.set_ku_tile(*best, precision=precision)
    store((sp.nx, sp.ny, sp.nz), precision, best, timings, P.device, persist)
    return best, timings


# ------------------------------------------------------------ scheduler shapes
# The device scheduler runs one 256-lane workgroup shape for every slot and a
# sweep depth ni per slot. Its choice is tuned per mix of grid sizes: the
# candidate shapes are timed on the batch itself (a few OC iterations of every
# job), first the tile with one depth for all grids, then the depth of each
# grid size in turn (largest grids first), keeping the others fixed.
SCHED_NIS = (4, 6, 8, 12, 16)
SCHED_START_NI = 8


def _grid_tag(g) -> str:
    return "x".join(str(int(v)) for v in g)


def sched_key(grids, device=None) -> str:
    gs = sorted({tuple(int(v) for v in g) for g in grids})
    return f"{_device_key(device)}|sched|" + ",".join(_grid_tag(g) for g in gs)


def lookup_sched(grids, device=None):
    """Cached (tile, {grid: ni}) for this mix of grid sizes, or None."""
    with _lock:
        v = _load().get(sched_key(grids, device))
    if not v:
        return None
    nis = {tuple(int(s) for s in k.split("x")): int(n) for k, n in v["ni"].items()}
    return tuple(v["tile"]), nis


def _pcg_bytes(P) -> int:
    """Algorithmic bytes of one minimal fused FP32 PCG iteration of P."""
    return 120 * P.spec.n_nodes + 4 * P.spec.n_elem


def time_sched(problems, tile, nis, n_iters=4, graph_ticks=32, **job) -> float:
    """Seconds per byte of algorithmic PCG traffic for one scheduler run of
    every problem (n_iters OC iterations each) with the given tile and
    per-grid depths. Normalising by the work done keeps candidates whose
    rounding changes the PCG iteration counts comparable."""
    from .sched import Scheduler
    dims = [(p.spec.nx, p.spec.ny, p.spec.nz, nis[(p.spec.nx, p.spec.ny, p.spec.nz)])
            for p in problems]
    S = Scheduler(dims, tile=tile, device=problems[0].device, graph_ticks=graph_ticks)
    try:
        for i, p in enumerate(problems):
            S.admit(i, p, n_iters=n_iters, **job)
        torch.cuda.synchronize(S.device)
        t0 = time.perf_counter()
        _, st = S.run()
        torch.cuda.synchronize(S.device)
        dt = time.perf_counter() - t0
    finally:
        torch.cuda.synchronize(S.device)
        S.close()
    work = sum(s.pcg_iters * _pcg_bytes(p) for s, p in zip(st, problems))
    return dt / max(work, 1)


def tune_sched(problems, n_iters=4, tiles=None, nis=SCHED_NIS, persist=True, log=None, **job):
    """Picks the scheduler's workgroup shape and per-grid sweep depths for
    this batch and caches them. Returns (tile, {grid: ni}, timings)."""
    from .sched import TILES as SCHED_TILES
    tiles = list(tiles or SCHED_TILES)
    grids = sorted({(p.spec.nx, p.spec.ny, p.spec.nz) for p in problems},
                   key=lambda g: -g[0] * g[1] * g[2])
    timings = {}

    def trial(tile, depth):
        k = (tuple(tile), tuple(depth[g] for g in grids))
        if k not in timings:
            timings[k] = time_sched(problems, tile, depth, n_iters, **job)
            if log:
                log(f"[autotune] sched tile={tile} ni={dict((_grid_tag(g), depth[g]) for g in grids)} "
                    f"{timings[k] * 1e12:.2f} us/MB")
        return timings[k]

    depth = {g: SCHED_START_NI for g in grids}
    best_tile = min(tiles, key=lambda t: trial(t, depth))
    for g in grids:
        best_ni, best_s = depth[g], trial(best_tile, depth)
        for ni in nis:
            d = dict(depth)
            d[g] = ni
            s = trial(best_tile, d)
            if s < best_s:
                best_ni, best_s = ni, s
        depth[g] = best_ni
    with _lock:
        d = _load()
        d[sched_key(grids, problems[0].device)] = {
            "tile": list(best_tile), "ni": {_grid_tag(g): depth[g] for g in grids},
            "us_per_mb": trial(best_tile, depth) * 1e12, "candidates": len(timings),
            "oc_iters_timed": n_iters}
        if persist:
            CACHE.write_text(json.dumps(d, indent=1, sort_keys=True))
    return best_tile, depth, timings


def sched_shape(problems, n_iters=4, log=None):
    """The cached scheduler shape for this batch's grid sizes, tuned first
    when there is none: (tile, [ni of each problem])."""
    grids = [(p.spec.nx, p.spec.ny, p.spec.nz) for p in problems]
    hit = lookup_sched(grids, problems[0].device)
    if hit is None:
        tile, depth, _ = tune_sched(problems, n_iters=n_iters, log=log)
    else:
        tile, depth = hit
    return tile, [depth[g] for g in grids]


def _main(argv=None):
    """Tunes the single-problem tiles of the given grids (fp32 and fp64) and,
    with --batch48, the scheduler shape of the seed-0 48-problem batch."""
    import argparse
    import sys
    ap = argparse.ArgumentParser(prog="python -m genpipeline_amd.autotune")
    ap.add_argument("--grid", type=int, nargs=3, action="append", default=[])
    ap.add_argument("--batch48", action="store_true")
    ap.add_argument("--sched-oc-iters", type=int, default=4)
    a = ap.parse_args(argv)
    from .bcs import clamp_x0
    from .problem import Problem, ProblemSpec
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    dev = torch.device("cuda", 0)
    for nx, ny, nz in a.grid:
        sp = ProblemSpec(nx, ny, nz)
        f = np.zeros(sp.n_dof)
        f[-1] = -1.0
        with Problem(sp, clamp_x0(nx, ny, nz), f, device=dev) as P:
            for prec in ("fp32", "fp64"):
                best, t = tune(P, prec, log=log)
                print(f"{nx}x{ny}x{nz} {prec}: tile {best} {t[best]:.2f} us/iter", flush=True)
    if a.batch48:
        from .problems import batch_specs
        ws = batch_specs(seed=0, n=48)
        Ps = [Problem(w.spec, w.fixed, w.f, device=dev) for w in ws]
        try:
            tile, depth, t = tune_sched(Ps, n_iters=a.sched_oc_iters, log=log, max_iter=20000)
            print(f"batch48 scheduler: tile {tile} ni "
                  f"{ {_grid_tag(g): n for g, n in depth.items()} } ({len(t)} candidates)",
                  flush=True)
        finally:
            for P in Ps:
                P.close()


if __name__ == "__main__":
    _main()
