"""Many independent SIMP problems in one batched execution (``gp_batch``).

Every phase of the SIMP loop -- material law, Jacobi setup, each PCG
iteration, sensitivity, filters, every OC bisection step -- is one kernel
launch over all problems of the batch, whatever their grid sizes. A problem's
results do not depend on what else is in the batch: reductions are
per-problem and fixed-order, so a problem gives bitwise the same answer alone
(``Problem.run`` with the same K.u tile) and in any mix.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np
import torch

from . import _lib as L
from .problem import Problem, SimpResult, _stream

DEFAULT_TILE = (8, 32)


class Batch:
    """A fixed set of ``Problem`` objects run together.

    tile: (tj, tk) modal K.u workgroup shape shared by the batch; each problem
    keeps its own sweep depth (``Problem.set_ku_tile``'s ni).
    """

    def __init__(self, problems: Sequence[Problem], tile=None):
        if not problems:
            raise ValueError("empty batch")
        self.problems = list(problems)
        dev = {p.device for p in self.problems}
        if len(dev) != 1:
            raise ValueError("all problems of a batch must live on one device")
        self.device = self.problems[0].device
        tj, tk = tile or DEFAULT_TILE
        for p in self.problems:
            cur = p.ku_tile("fp32")
            ni = cur[2] if cur[0] else 2
            p.set_ku_tile(tj, tk, ni, precision="fp32")
        self.tile = (tj, tk)
        arr = (C.c_void_p * len(self.problems))(*[p.handle.value for p in self.problems])
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            L.call("gp_batch_create", arr, len(self.problems), tj, tk, C.byref(h))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.load().gp_batch_destroy(self._h)
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

    def run(self, n_iters, volfracs=None, tol=1e-6, max_iter=20000, warm_start=True, x0=None):
        """One SIMP run per problem; returns a list of SimpResult."""
        n = len(self.problems)
        vfs = [p.spec.volfrac for p in self.problems] if volfracs is None else list(volfracs)
        xs, xps, us = [], [], []
        for b, p in enumerate(self.problems):
            sp = p.spec
            if x0 is not None and x0[b] is not None:
                x = torch.as_tensor(x0[b], dtype=torch.float64, device=self.device).clone()
            else:
                x = torch.full((sp.n_elem,), float(vfs[b]), dtype=torch.float64, device=self.device)
                if p.preserved is not None:
                    x[torch.from_numpy(p.preserved).to(self.device)] = 1.0
            xs.append(x)
            xps.append(x.clone())
            us.append(torch.zeros(sp.n_dof, dtype=torch.float64, device=self.device))
        o = L.GpSimpOpts()
        o.n_iters = int(n_iters)
        o.pcg.precision = L.GP_F32
        o.pcg.max_iter = int(max_iter)
        o.pcg.tol = float(tol)
        o.warm_start = 1 if warm_start else 0
        vf_arr = (C.c_double * n)(*vfs)
        px = (C.c_void_p * n)(*[t.data_ptr() for t in xs])
        pxp = (C.c_void_p * n)(*[t.data_ptr() for t in xps])
        pu = (C.c_void_p * n)(*[t.data_ptr() for t in us])
        st = (L.GpSimpStats * n)()
        L.call("gp_batch_simp_run", self._h, C.byref(o), vf_arr, px, pxp, pu, st, _stream(self.device))
        return [SimpResult(xs[b], xps[b], us[b], st[b].compliance, st[b].iters_done,
                           st[b].pcg_iters_total, st[b].bisect_steps_total, st[b].status,
                           st[b].seconds) for b in range(n)]
