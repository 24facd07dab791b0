"""Continuous-batching request loop over the device scheduler.

Requests -- density voxel fields with a kind -- arrive over time (open loop).
A request is admitted as soon as a scheduler slot of its grid size is free
and its problem fits the HBM budget; it then runs in the shared ticks beside
everything else on the device, and its slot is released and refilled the
moment it finishes. State (density, displacement, PCG vectors) stays on the
device from admission to completion.

Request kinds mirror the reference entry points the BO loop issues:

* ``"solve"`` -- one linear-elastic solve with GPUVoxelFEM semantics
  (genpipeline/fem/gpu_fem_solver.py:76-133): E = max(voxel, 1e-3), supports
  and load of SIMPSolverGPU._get_bcs with force_n, outputs compliance, the
  stress proxy 10 C / n_solid, the displacement proxy, and the sentinels
  {1e4, 1e3, 1e6} when fewer than 50 voxels are solid or C < 1. Solved by
  FP64 iterative refinement (to 1e-8 by default) around FP32 inner solves.
* ``"simp"`` -- SIMPSolverGPU.run(volfrac, n_iters, x_init=density)
  (genpipeline/topology/simp_solver_gpu.py:71-136): x_init rescaled to the
  volume fraction and clipped to [1e-3, 1], tolerance schedule, compliance
  solve after every update, early stop; outputs xPhys and compliance.

A request's results are bitwise the same alone and in any mix (the
scheduler's per-problem reductions are fixed-order and independent of the
other slots).
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from .bcs import simp_gpu_bcs
from .problem import Problem, ProblemSpec, footprint_bytes
from .problems import Request as ServeRequest  # rid, t_arrive, shape, density, kind, volfrac
from .sched import Scheduler

KINDS = ("solve", "simp")


@dataclass
class ServeResult:
    rid: int
    kind: str
    shape: tuple
    compliance: float
    stress_max: Optional[float]
    displacement_max: Optional[float]
    iters_done: int
    pcg_iters: int
    status: int
    t_arrive: float
    t_admit: float
    t_done: float
    xPhys: Optional[torch.Tensor] = field(default=None, repr=False)
    u: Optional[torch.Tensor] = field(default=None, repr=False)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrive


@dataclass
class ServeConfig:
    force_n: float = 1000.0       # "solve" total load (GPUVoxelFEM default)
    simp_force: float = 1.0       # "simp" load magnitude (SIMPSolverGPU.run default)
    simp_iters: int = 30
    # "solve" requests: FP64 iterative refinement to solve_tol around FP32
    # inner solves at solve_inner_tol (refine False: one FP32 solve to
    # solve_tol, whose accuracy stops near 1e-5 on solid/void fields)
    solve_refine: bool = True
    solve_tol: float = 1e-8
    solve_inner_tol: float = 1e-3
    solve_max_outer: int = 50
    # "simp" requests: the compliance solve after each update starts from the
    # previous displacement instead of zero (SIMPSolverGPU solves it from zero,
    # simp_solver_gpu.py:747-751; same answers to the tolerance, fewer iterations)
    simp_warm_compliance: bool = False
    # "simp" requests: each warm OC solve starts from the Galerkin combination
    # of the request's last simp_recycle solutions (0: from the last u, as
    # SIMPSolverGPU does); the ring adds 8 * simp_recycle bytes per DOF
    simp_recycle: int = 0
    max_iter: int = 20000
    # admission order among queued requests that fit: "fifo" (arrival order)
    # or "sjf" (single solves before SIMP runs, each kind in arrival order;
    # a SIMP run costs about twenty solves)
    policy: str = "fifo"
    bcs: dict = field(default_factory=lambda: {"fixed_face": "x_min", "load_face": "x_max",
                                               "load_dof": 2})


def _spec(shape, kind) -> ProblemSpec:
    nx, ny, nz = shape
    if kind == "solve":
        return ProblemSpec(nx, ny, nz, penal=1.0, law="clamp", oc="gpu", filter_weights="f32")
    return ProblemSpec(nx, ny, nz, penal=3.0, law="clamp", oc="gpu", filter_weights="f32")


def _flat64(density):
    """A request's density as a flat float64 array (a device tensor stays on
    its device)."""
    if isinstance(density, torch.Tensor):
        return density.reshape(-1).to(torch.float64)
    return np.asarray(density, np.float64).ravel()


def job_args(req: ServeRequest, cfg: ServeConfig) -> dict:
    """Scheduler job arguments of a request (also used to run it alone)."""
    dens = _flat64(req.density)
    if req.kind == "solve":
        tol = cfg.solve_tol if getattr(req, "tol", None) is None else float(req.tol)
        refine = cfg.solve_refine and getattr(req, "refine", True)
        extra = (dict(refine=True, inner_tol=cfg.solve_inner_tol, max_outer=cfg.solve_max_outer)
                 if refine else {})
        return dict(n_iters=0, x0=dens, xphys0=dens, tol=tol, max_iter=cfg.max_iter,
                    warm_start=False, rhs=getattr(req, "rhs", None), **extra)
    if isinstance(dens, torch.Tensor):
        dens = dens.cpu().numpy()
    x = np.clip(dens * (req.volfrac / max(dens.mean(), 1e-6)), 1e-3, 1.0)
    return dict(n_iters=cfg.simp_iters, volfrac=req.volfrac, x0=x, xphys0=x, tol=1e-6,
                max_iter=cfg.max_iter, tol_schedule=True, early_stop=True,
                compliance_each_iter=True, warm_start=True,
                warm_compliance=cfg.simp_warm_compliance, recycle=cfg.simp_recycle)


def problem_bytes(shape, kind, cfg: ServeConfig) -> int:
    """Device bytes a pooled problem of this grid and kind holds, with the
    recycled-start ring of a "simp" problem."""
    return footprint_bytes(_spec(shape, kind),
                           recycle=cfg.simp_recycle if kind == "simp" else 0)


def make_problem(shape, kind, cfg: ServeConfig, device) -> Problem:
    spec = _spec(shape, kind)
    force = cfg.force_n if kind == "solve" else cfg.simp_force
    f, fixed = simp_gpu_bcs(cfg.bcs, *shape, force)
    if kind == "solve":  # GPUVoxelFEM builds its load in float32 (gpu_fem_solver.py:86)
        f = f.astype(np.float32).astype(np.float64)
    return Problem(spec, fixed, f, device=device)


def _fem_rule(compliance: float, n_solid: int, disp: float):
    """GPUVoxelFEM's output proxies and sentinels (gpu_fem_solver.py:96-133)."""
    if n_solid < 50 or compliance < 1.0:
        return 1e6, 1e4, 1e3
    return (float(np.clip(compliance, 0, 1e8)),
            float(np.clip(compliance / max(n_solid, 1) * 10.0, 0, 1e4)),
            float(np.clip(disp * 1e-3 * 1e3, 0, 1e4)))


def fem_outputs(compliance: float, u: torch.Tensor, density: np.ndarray):
    """(compliance, stress_max, displacement_max) with GPUVoxelFEM's proxies and
    sentinels (gpu_fem_solver.py:96-133)."""
    disp = float(u.view(-1, 3).norm(dim=1).max().item())
    if isinstance(density, torch.Tensor):
        n_solid = int((density > 0.5).sum().item())
    else:
        n_solid = int((np.asarray(density) > 0.5).sum())
    return _fem_rule(compliance, n_solid, disp)


def fem_outputs_batch(compliances, us, densities: torch.Tensor) -> list:
    """fem_outputs of B fields of one grid (densities: a (B, ...) device
    tensor) with two read-backs for the batch instead of two per field."""
    if not len(us):
        return []
    disp = torch.stack([u.view(-1, 3) for u in us]).norm(dim=2).amax(dim=1).tolist()
    n_solid = (densities.reshape(len(us), -1) > 0.5).sum(dim=1).tolist()
    return [_fem_rule(c, int(n), d) for c, n, d in zip(compliances, n_solid, disp)]


class Server:
    """slots: {shape: number of slots}; hbm_budget: bytes of problem state
    the server may hold (default: half the free device memory)."""

    def __init__(self, slots: dict, cfg: ServeConfig | None = None, hbm_budget: int | None = None,
                 tile=(16, 16), ni: int = 4, graph_ticks: int = 32, chunk_ticks: int = 64,
                 device=None):
        if cfg is not None and cfg.policy not in ("fifo", "sjf"):
            raise ValueError(f"unknown admission policy {cfg.policy!r}")
        if not torch.cuda.is_available():
            raise RuntimeError("Server needs a ROCm GPU (torch.cuda.is_available() is False)")
        d = torch.device("cuda") if device is None else torch.device(device)
        self.device = d if d.index is not None else torch.device("cuda", torch.cuda.current_device())
        self.cfg = cfg or ServeConfig()
        dims, self.slot_shape = [], []
        for shape, n in slots.items():
            for _ in range(int(n)):
                dims.append((*shape, ni))
                self.slot_shape.append(tuple(shape))
        self.S = Scheduler(dims, tile=tile, device=self.device, graph_ticks=graph_ticks)
        self.chunk = int(chunk_ticks)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = free // 2
        self.budget = int(hbm_budget)
        self.used = 0
        self.pool: dict = {}
        self.peak_used = 0

    def close(self):
        self.S.close()
        for ps in self.pool.values():
            for p in ps:
                p.close()
        self.pool = {}

    def _take_problem(self, shape, kind) -> Optional[Problem]:
        ps = self.pool.setdefault((shape, kind), [])
        if ps:
            return ps.pop()
        fp = problem_bytes(shape, kind, self.cfg)
        # make room by freeing idle problems of other kinds or sizes
        for key, idle in self.pool.items():
            while idle and self.used + fp > self.budget:
                q = idle.pop()
                self.used -= problem_bytes(*key, self.cfg)
                q.close()
        if self.used + fp > self.budget:
            return None
        self.used += fp
        self.peak_used = max(self.peak_used, self.used)
        return make_problem(shape, kind, self.cfg, self.device)

    def run(self, requests: Sequence[ServeRequest], time_scale: float = 1.0,
            keep_xphys: bool = False, keep_u: bool = False,
            outputs: bool = True) -> list[ServeResult]:
        """Serve requests arriving at t_arrive * time_scale seconds after the
        call; returns one result per request (request order). keep_xphys /
        keep_u keep each job's density / displacement (device tensors).
        outputs=False skips the FEM output formulas of finished solves (each
        reads two device reductions back): compliance is then the solver's,
        stress_max and displacement_max None."""
        reqs = sorted(requests, key=lambda r: r.t_arrive)
        results: dict[int, ServeResult] = {}
        queue: deque = deque()
        active: dict[int, tuple] = {}  # slot -> (request, problem, t_admit)
        free = {}
        for s, shp in enumerate(self.slot_shape):
            free.setdefault(shp, []).append(s)
        for r in reqs:
            if tuple(r.shape) not in free:
                raise ValueError(f"no slot for grid {r.shape}")
        nxt = 0
        t0 = time.perf_counter()
        while len(results) < len(reqs):
            now = time.perf_counter() - t0
            while nxt < len(reqs) and reqs[nxt].t_arrive * time_scale <= now:
                queue.append(reqs[nxt])
                nxt += 1
            # admission: in policy order among requests that fit
            order = list(queue)
            if self.cfg.policy == "sjf":
                order.sort(key=lambda q: (q.kind != "solve", q.t_arrive))
            for r in order:
                shp = tuple(r.shape)
                if not free[shp]:
                    continue
                P = self._take_problem(shp, r.kind)
                if P is None:
                    continue
                slot = free[shp].pop()
                self.S.admit(slot, P, **job_args(r, self.cfg))
                active[slot] = (r, P, time.perf_counter() - t0)
                queue.remove(r)
            if not active:
                if queue:
                    raise RuntimeError(f"request {queue[0].rid} ({queue[0].shape}) does not fit "
                                       f"the HBM budget of {self.budget} bytes")
                if nxt < len(reqs):
                    wait = reqs[nxt].t_arrive * time_scale - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.01))
                continue
            self.S.tick(self.chunk)
            st = self.S.status()
            t_now = time.perf_counter() - t0
            for slot in [s for s in active if st[s].done]:
                r, P, t_admit = active.pop(slot)
                job = self.S.jobs[slot]
                s = st[slot]
                if r.kind == "solve" and outputs:
                    c, sm, dm = fem_outputs(s.compliance, job.u, r.density)
                else:
                    c, sm, dm = s.compliance, None, None
                results[r.rid] = ServeResult(r.rid, r.kind, tuple(r.shape), c, sm, dm,
                                             s.iters_done, s.pcg_iters, s.status, r.t_arrive * time_scale,
                                             t_admit, t_now,
                                             job.xPhys if keep_xphys else None,
                                             job.u if keep_u else None)
                self.S.release(slot)
                free[tuple(r.shape)].append(slot)
                self.pool[(tuple(r.shape), r.kind)].append(P)
        return [results[r.rid] for r in requests]


def run_alone(req: ServeRequest, cfg: ServeConfig | None = None, tile=(16, 16), ni: int = 4,
              device=None) -> ServeResult:
    """The same request on an otherwise empty scheduler (reference for the
    alone-versus-mix check)."""
    cfg = cfg or ServeConfig()
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    P = make_problem(tuple(req.shape), req.kind, cfg, dev)
    S = Scheduler([(*req.shape, ni)], tile=tile, device=dev)
    try:
        S.admit(0, P, **job_args(req, cfg))
        _, st = S.run()
        job, s = S.jobs[0], st[0]
        if req.kind == "solve":
            c, sm, dm = fem_outputs(s.compliance, job.u, req.density)
        else:
            c, sm, dm = s.compliance, None, None
        return ServeResult(req.rid, req.kind, tuple(req.shape), c, sm, dm, s.iters_done,
                           s.pcg_iters, s.status, 0.0, 0.0, 0.0, job.xPhys)
    finally:
        torch.cuda.synchronize(dev)
        S.close()
        P.close()


def latency_summary(results: Sequence[ServeResult], wall: float) -> dict:
    lat = np.array([r.latency for r in results])
    return {"n": len(results), "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "mean_ms": float(lat.mean() * 1e3),
            "solves_per_s": len(results) / wall if wall > 0 else float("nan")}
