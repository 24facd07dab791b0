"""Device-side scheduler (``gp_sched``): independent problems in shared ticks.

A scheduler owns ``n`` slots. Each slot runs one job -- a SIMP run or a
single solve -- on one ``Problem``. One tick launch advances every occupied
slot by one phase of its job (material law, Jacobi set-up, one PCG iteration,
sensitivity, filter, one OC bisection step, commit, compliance), so slots
progress independently: small problems finish early, and a finished slot can
be released and refilled between ticks while the others keep running
(continuous batching). Ticks can be replayed from a captured HIP graph.

A job's results are bitwise those of ``Problem.run`` with the same K.u tile
(tj, tk, ni), whatever else runs in the other slots.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from .problem import Problem, SimpResult, _stream

TILES = ((8, 32), (16, 16), (14, 18), (6, 34), (7, 34))


@dataclass
class SlotStatus:
    phase: int
    done: bool
    iters_done: int
    pcg_iters: int
    bisect_steps: int
    status: int
    compliance: float


@dataclass
class Job:
    problem: Problem
    x: torch.Tensor
    xPhys: torch.Tensor
    u: torch.Tensor
    n_iters: int
    volfrac: float
    rhs: Optional[torch.Tensor] = None


def _status(s: L.GpSlotStatus) -> SlotStatus:
    return SlotStatus(s.phase, bool(s.done), s.iters_done, s.pcg_iters_total,
                      s.bisect_steps_total, s.status, s.compliance)


class Scheduler:
    """slot_dims: per slot (nx, ny, nz, ni): the largest grid the slot takes
    and the modal sweep depth of its jobs. tile: (tj, tk) in TILES.
    graph_ticks: even number of ticks captured in one HIP graph (0: none)."""

    def __init__(self, slot_dims: Sequence[Sequence[int]], tile=(8, 32), device=None,
                 graph_ticks: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("Scheduler needs a ROCm GPU (torch.cuda.is_available() is False)")
        if tuple(tile) not in TILES:
            raise ValueError(f"tile must be one of {TILES}")
        d = torch.device("cuda") if device is None else torch.device(device)
        self.device = d if d.index is not None else torch.device("cuda", torch.cuda.current_device())
        self.tile = tuple(tile)
        self.slot_dims = [tuple(int(v) for v in sd) for sd in slot_dims]
        flat = (C.c_int32 * (4 * len(self.slot_dims)))(*[v for sd in self.slot_dims for v in sd])
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            L.call("gp_sched_create", len(self.slot_dims), flat, self.tile[0], self.tile[1],
                   C.byref(h))
        self._h = h
        self.jobs: list[Optional[Job]] = [None] * len(self.slot_dims)
        self.graph_ticks = 0
        if graph_ticks:
            self.set_graph(graph_ticks)

    @property
    def n_slots(self) -> int:
        return len(self.slot_dims)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.load().gp_sched_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def blocks(self, slot: int = -1) -> int:
        out = C.c_int32()
        L.call("gp_sched_blocks", self._h, int(slot), C.byref(out))
        return out.value

    def set_graph(self, ticks: int):
        with torch.cuda.device(self.device):
            L.call("gp_sched_set_graph", self._h, int(ticks))
        self.graph_ticks = int(ticks)

    def admit(self, slot: int, problem: Problem, n_iters: int = 0, volfrac: float | None = None,
              tol: float = 1e-6, max_iter: int = 20000, tol_schedule: bool = False,
              early_stop: bool = False, compliance_each_iter: bool = False,
              warm_start: bool = True, x0=None, xphys0=None, u0=None, rhs=None,
              refine: bool = False, inner_tol: float = 1e-3, max_outer: int = 50,
              warm_compliance: bool = False, recycle: int = 0) -> Job:
        """Start a job in a free slot (takes effect at the next tick).

        n_iters = 0 runs one solve at the given density (``xphys0``, else
        ``x0``, else volfrac) and reports its compliance. ``rhs`` (n_dof,
        single solves only) replaces the problem's load: the job solves
        K d = rhs from d = 0 (``warm_start`` must be False). ``refine``: every
        solve is FP64 iterative refinement around FP32 inner solves at
        ``inner_tol`` (``tol`` the FP64 residual tolerance), bitwise
        ``Problem.run(refine=True)``. ``warm_compliance``: compliance solves
        start from the previous u (as ``Problem.run``'s option). ``recycle``
        = m: warm solves start from the Galerkin combination of the job's last
        m solutions, bitwise ``Problem.run(recycle=m)``; the ring is the
        problem's own, so a problem runs in one such job at a time."""
        sp = problem.spec
        vf = sp.volfrac if volfrac is None else float(volfrac)
        dev = self.device
        if x0 is None:
            x = torch.full((sp.n_elem,), vf, dtype=torch.float64, device=dev)
            if problem.preserved is not None:
                x[torch.from_numpy(problem.preserved).to(dev)] = 1.0
        else:
            x = torch.as_tensor(x0, dtype=torch.float64, device=dev).reshape(-1).clone()
        xp = x.clone() if xphys0 is None else torch.as_tensor(
            xphys0, dtype=torch.float64, device=dev).reshape(-1).clone()
        u = (torch.zeros(sp.n_dof, dtype=torch.float64, device=dev) if u0 is None
             else torch.as_tensor(u0, dtype=torch.float64, device=dev).reshape(-1).clone())
        if x.numel() != sp.n_elem or xp.numel() != sp.n_elem or u.numel() != sp.n_dof:
            raise ValueError("state shapes do not match the problem")
        # every argument check comes before the device slot is taken, so a
        # rejected call leaves the slot free
        f = None
        if rhs is not None:
            if n_iters != 0 or warm_start:
                raise ValueError("rhs applies to single solves from zero (n_iters=0, "
                                 "warm_start=False)")
            f = torch.as_tensor(rhs, dtype=torch.float64, device=dev).reshape(-1).contiguous()
            if f.numel() != sp.n_dof:
                raise ValueError("rhs length does not match the problem")
        j = L.GpJob()
        j.n_iters = int(n_iters)
        j.volfrac = vf
        j.tol = float(tol)
        j.max_iter = int(max_iter)
        j.tol_schedule = 1 if tol_schedule else 0
        j.early_stop = 1 if early_stop else 0
        j.compliance_each_iter = 1 if compliance_each_iter else 0
        j.warm_start = 1 if warm_start else 0
        j.refine = 1 if refine else 0
        j.inner_tol = float(inner_tol)
        j.max_outer = int(max_outer)
        j.warm_compliance = 1 if warm_compliance else 0
        j.recycle = int(recycle)
        with torch.cuda.device(dev):
            L.call("gp_sched_admit", self._h, int(slot), problem.handle, C.byref(j),
                   C.c_void_p(x.data_ptr()), C.c_void_p(xp.data_ptr()), C.c_void_p(u.data_ptr()),
                   _stream(dev))
        if f is not None:
            try:
                with torch.cuda.device(dev):
                    L.call("gp_sched_set_rhs", self._h, int(slot), C.c_void_p(f.data_ptr()),
                           _stream(dev))
            except Exception:
                # the slot holds pointers to this call's tensors: free it
                # before they go
                with torch.cuda.device(dev):
                    L.call("gp_sched_release", self._h, int(slot), _stream(dev))
                raise
        job = Job(problem, x, xp, u, int(n_iters), vf, f)
        self.jobs[slot] = job
        return job

    def release(self, slot: int):
        with torch.cuda.device(self.device):
            L.call("gp_sched_release", self._h, int(slot), _stream(self.device))
        self.jobs[slot] = None

    def tick(self, n: int = 1):
        with torch.cuda.device(self.device):
            L.call("gp_sched_tick", self._h, int(n), _stream(self.device))

    def status(self) -> list[SlotStatus]:
        arr = (L.GpSlotStatus * self.n_slots)()
        with torch.cuda.device(self.device):
            L.call("gp_sched_status", self._h, arr, _stream(self.device))
        return [_status(s) for s in arr]

    def run(self, max_ticks: int = 1 << 40):
        """Tick until every occupied slot is done; returns (ticks, statuses)."""
        arr = (L.GpSlotStatus * self.n_slots)()
        ticks = C.c_int64(0)
        with torch.cuda.device(self.device):
            L.call("gp_sched_run", self._h, int(max_ticks), C.byref(ticks), arr,
                   _stream(self.device))
        return ticks.value, [_status(s) for s in arr]

    def result(self, slot: int, st: SlotStatus) -> SimpResult:
        j = self.jobs[slot]
        return SimpResult(j.x, j.xPhys, j.u, st.compliance, st.iters_done, st.pcg_iters,
                          st.bisect_steps, st.status, 0.0)


def slot_dims_for(problems: Sequence[Problem], ni: int | None = None):
    """Exact-capacity slots, one per problem; ni defaults to each problem's
    tuned sweep depth (2 when it has none)."""
    out = []
    for p in problems:
        t = p.ku_tile("fp32")
        d = ni if ni is not None else (t[2] if t[0] else 2)
        out.append((p.spec.nx, p.spec.ny, p.spec.nz, max(1, int(d))))
    return out


def run_batch(problems: Sequence[Problem], n_iters: int, volfracs=None, tile=(8, 32), ni=None,
              graph_ticks: int = 32, **job) -> tuple[list[SimpResult], int]:
    """One job per problem, all slots admitted at once, run to completion.
    Returns (results, ticks). tile="auto" takes the workgroup shape and the
    per-grid sweep depths the autotuner picked for this mix of grid sizes
    (tuning it first when none is cached); ni is then ignored."""
    if isinstance(tile, str):
        if tile != "auto":
            raise ValueError(f"tile must be one of {TILES} or 'auto'")
        from . import autotune
        tile, nis = autotune.sched_shape(problems)
        dims = [(p.spec.nx, p.spec.ny, p.spec.nz, n) for p, n in zip(problems, nis)]
    else:
        dims = slot_dims_for(problems, ni)
    S = Scheduler(dims, tile=tile, device=problems[0].device, graph_ticks=graph_ticks)
    try:
        vfs = [p.spec.volfrac for p in problems] if volfracs is None else list(volfracs)
        for i, p in enumerate(problems):
            S.admit(i, p, n_iters=n_iters, volfrac=vfs[i], **job)
        ticks, st = S.run()
        return [S.result(i, st[i]) for i in range(len(problems))], ticks
    finally:
        torch.cuda.synchronize(S.device)
        S.close()
