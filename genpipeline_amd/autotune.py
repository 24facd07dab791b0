"""Autotuner for the modal K.u workgroup shape.

For a grid size and precision it times a fixed number of PCG iterations
(tolerance 0, so every candidate does the same work) for each supported
(tj, tk) tile and a range of sweep depths ni, and keeps the fastest. Results
are cached per (device ISA, grid, precision) in ``tuned_tiles.json`` next to
this file; ``Problem`` consults the cache when it is created.
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
NIS = [1, 2, 3, 4, 6, 8, 12, 16, 24, 32]
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


def time_pcg(P, E, tile, precision="fp32", iters=60, repeats=2) -> float:
    """Microseconds per PCG iteration with the given tile (tol = 0)."""
    P.set_ku_tile(*tile, precision=precision)
    P.pcg(E, tol=0.0, max_iter=4, precision=precision)  # warm the code path
    best = float("inf")
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, r = P.pcg(E, tol=0.0, max_iter=iters, precision=precision)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / max(r.iters, 1)
        best = min(best, dt)
    return best * 1e6


def tune(P, precision="fp32", cands=None, iters=60, repeats=2, persist=True, log=None):
    """Times every candidate on problem P and stores the fastest tile."""
    sp = P.spec
    rng = np.random.default_rng(0)
    E = torch.from_numpy(rng.uniform(1e-3, 1.0, sp.n_elem).astype(np.float32)).to(P.device)
    cands = cands or candidates(sp.nx)
    timings = {}
    for c in cands:
        timings[tuple(c)] = time_pcg(P, E, c, precision, iters, repeats)
        if log:
            log(f"[autotune] {sp.nx}x{sp.ny}x{sp.nz} {precision} tile={c} "
                f"{timings[tuple(c)]:.2f} us/iter")
    best = min(timings, key=timings.get)
    P.set_ku_tile(*best, precision=precision)
    store((sp.nx, sp.ny, sp.nz), precision, best, timings, P.device, persist)
    return best, timings
