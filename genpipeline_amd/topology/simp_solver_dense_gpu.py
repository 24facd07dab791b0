"""DenseGPUSolver: the reference "dense" GPU SIMP solver's interface.

Drop-in for genpipeline/topology/simp_solver_dense_gpu.py:20-391 (constructor
and ``run``), routed to the matrix-free hex8 path. The reference builds a
dense n_dof x n_dof matrix (3.3 GB at 32x16x16 in FP32) from a placeholder
element matrix (:104-138) with its own corner order (:160-168); none of that
is reproducible physics, so only the problem set-up is kept:

  * supports: every DOF of the nodes on faces x_min, y_min and z_min
    (:221-226);
  * load: -force_mag / (number of x_max nodes) in z on every x_max node, the
    share of x_max nodes that are also supported being dropped (:228-235 and
    the boundary-row zeroing of :251-253);
  * E = max(x, 1e-3)^penal (the reference uses x^3 with a 1e-5 diagonal
    shift, :33, :258), OC with l1 = 0, l2 = 1e9, <= 50 steps, gap 1e-4
    (:366-391), no preserved voxels;
  * Jacobi-PCG to relative residual 1e-5 from a zero start (:263-300);
  * ``last_compliance`` = f.u of the last solve, taken before the final OC
    update (:281-282).

There is no parity target for this solver; its tests check these semantics
against the oracle driven with the same set-up.
"""
from __future__ import annotations

import numpy as np
import torch

from ._common import cached_problem


def dense_bcs(nx: int, ny: int, nz: int, force_mag: float):
    """(f, fixed_dofs) of simp_solver_dense_gpu.py:209-235 (host, float64)."""
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    nid = i * (ny + 1) * (nz + 1) + j * (nz + 1) + k
    fixed_nodes = nid[(i == 0) | (j == 0) | (k == 0)]
    fixed = np.concatenate([3 * fixed_nodes, 3 * fixed_nodes + 1, 3 * fixed_nodes + 2]).astype(np.int64)
    load_nodes = nid[i == nx]
    f = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    f[3 * load_nodes + 2] = -float(force_mag) / len(load_nodes)
    f[fixed] = 0.0
    return f, fixed


class DenseGPUSolver:
    """simp_solver_dense_gpu.py:23-31 constructor; ``tol``/``max_iter`` are
    keyword-only extras (the reference's 1e-5 and a Jacobi-sized cap)."""

    def __init__(self, nx: int = 32, ny: int = 8, nz: int = 8, penal: float = 3.0,
                 rmin: float = 1.5, device: str = "cuda", *, tol: float = 1e-5,
                 max_iter: int = 20000):
        if not torch.cuda.is_available():
            raise RuntimeError("DenseGPUSolver needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.penal = float(penal)
        self.rmin = float(rmin)
        self.device = device
        self._torch_device = torch.device(device) if str(device) != "cuda" else None
        self.n_elem = self.nx * self.ny * self.nz
        self.n_nodes = (self.nx + 1) * (self.ny + 1) * (self.nz + 1)
        self.n_dof = 3 * self.n_nodes
        self.tol = float(tol)
        self.max_iter = int(max_iter)
        self.last_compliance = 0.0

    def _get_bcs(self, force_mag: float):
        f, fixed = dense_bcs(self.nx, self.ny, self.nz, force_mag)
        dev = self._torch_device or torch.device("cuda", torch.cuda.current_device())
        return torch.from_numpy(f).float().to(dev), torch.from_numpy(fixed).int().to(dev)

    def run(self, volfrac: float = 0.25, n_iters: int = 15, force_mag: float = 1000) -> np.ndarray:
        f, fixed = dense_bcs(self.nx, self.ny, self.nz, force_mag)
        P = cached_problem(self, "clamp", "cpu", f, fixed, None, volfrac)
        n_iters = int(n_iters)
        if n_iters <= 0:
            return np.full((self.nx, self.ny, self.nz), float(volfrac), dtype=np.float32)
        kw = dict(tol=self.tol, max_iter=self.max_iter, precision="fp32")
        # n_iters - 1 whole iterations, then the last one step by step so that
        # last_compliance is the solve before the final update, as in the reference
        r = P.run(n_iters - 1, volfrac=float(volfrac), warm_start=False, **kw)
        x, xPhys = r.x, r.xPhys
        E = P.stiffness(xPhys)
        u, _ = P.pcg(E, **kw)
        self.last_compliance = P.compliance(u)
        dc = P.filter(P.sens(xPhys, u), mode=1)
        _, xPhys, _, _ = P.oc(x, dc, volfrac=float(volfrac))
        return xPhys.float().cpu().numpy().reshape(self.nx, self.ny, self.nz)
