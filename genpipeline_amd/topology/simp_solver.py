"""SIMPSolver: the reference CPU SIMP solver's interface on the device path.

Drop-in for genpipeline/topology/simp_solver.py:13-79. Constructor, ``run``,
``last_compliance`` and the static ``_get_Ke`` keep the reference's names,
defaults and return types; the loop runs on the GPU through the C-ABI
library (``gp_simp_run``).

Semantics kept from the reference:
  * material law E = max(xPhys, 1e-3)^penal (simp_solver.py:201-203);
  * dc = -penal max(xPhys, 1e-3)^(penal-1) u_e^T KE u_e (:215-218);
  * density filter and its adjoint (:221-223), OC bisection with l1 = 0,
    l2 = 1e9, <= 50 steps, relative gap 1e-4, move 0.2, preserved voxels
    pinned to 1 in x and xPhys (:225-249);
  * x0 = volfrac (preserved = 1), xPhys0 = x0 unfiltered (:60-67);
  * boundary conditions of ``_get_bcs`` (:81-175) including every branch;
  * ``last_compliance`` = f.u of one more solve at the returned density (:78).

Differences (defaults; ``ke=`` and ``solver="direct"`` give the reference's
literal behaviour):
  * the element matrix is the exact trilinear hex8 matrix at nu = 0.3
    (2x2x2 Gauss); the reference's literal table (:284-1543) is neither
    symmetric nor positive semi-definite, so no iterative solver converges on
    it;
  * K u = f is solved by FP32 Jacobi-PCG with FP64 iterative refinement to a
    relative residual of ``tol`` (default 1e-10) instead of SuperLU; results
    agree with a direct solve to that tolerance.

Reference-literal mode: ``SIMPSolver(..., ke=table)`` with the reference's
own ``SIMPSolver._get_Ke(0.3)`` assembles that table as the reference does
(the np.kron index vectors of :280-281 put entry (a, b) of the table at
(edof[b], edof[a]), i.e. K = sum_e E_e KE^T), and ``solver="direct"`` (the
default whenever ``ke`` is given) solves every system exactly in FP64 on the
device by banded LU with partial pivoting (gp_direct), SuperLU's role
(:213); moduli, sensitivities, filter and OC are FP64 throughout.
"""
from __future__ import annotations

import numpy as np

from .. import _lib
from ..bcs import simp_bcs
from ._common import build_edof_mapping, cached_problem


class SIMPSolver:
    """3D SIMP topology optimisation on a structured voxel grid.

    Parameters follow simp_solver.py:24-54; keyword-only extras: ``device``
    (GPU to use), ``tol`` (relative residual of every PCG solve), ``ke``
    (24x24 element table, reference orientation; None: exact hex8) and
    ``solver`` ("pcg" or "direct"; default "direct" when ``ke`` is given).
    """

    def __init__(self, nx: int = 32, ny: int = 8, nz: int = 8, penal: float = 3.0,
                 rmin: float = 1.5, boundary_conditions: dict | None = None,
                 preserved_mask: np.ndarray | None = None, *, device=None, tol: float = 1e-10,
                 ke: np.ndarray | None = None, solver: str | None = None):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.penal = float(penal)
        self.rmin = float(rmin)
        if ke is None:
            self.Ke = self._get_Ke(nu=0.3)
            self._ke_dev = None
        else:
            self.Ke = np.array(ke, dtype=np.float64).reshape(24, 24)
            self._ke_dev = np.ascontiguousarray(self.Ke.T)  # K = sum_e E_e KE^T
        self.solver = solver or ("pcg" if ke is None else "direct")
        if self.solver not in ("pcg", "direct"):
            raise ValueError(f"solver must be 'pcg' or 'direct', not {self.solver!r}")
        self.last_compliance = 0.0
        self.preserved_mask = (None if preserved_mask is None
                               else np.asarray(preserved_mask).flatten().astype(bool))
        self.bcs = {"fixed_face": "x_min", "load_face": "x_max", "load_dof": 2}
        if boundary_conditions:
            self.bcs.update(boundary_conditions)
        self.device = device
        self._torch_device = device
        self.tol = float(tol)
        self.last_result = None

    # ------------------------------------------------------------ reference API
    @staticmethod
    def _get_Ke(nu: float = 0.3) -> np.ndarray:
        """24x24 element stiffness (E = 1, unit cube, C3D8 node order)."""
        return _lib.hex8_ke(float(nu))

    def _build_edof_mapping(self):
        """(edof_mat, iK, jK) with the reference's layout (simp_solver.py:251-282);
        the device path derives these indices arithmetically and never reads
        them."""
        return build_edof_mapping(self.nx, self.ny, self.nz)

    def _get_bcs(self, force_mag: float):
        """(f, fixed_dofs) as numpy arrays (simp_solver.py:81-157)."""
        return simp_bcs(self.bcs, self.nx, self.ny, self.nz, force_mag)

    def run(self, volfrac: float = 0.4, n_iters: int = 80, force_mag: float = 1.0) -> np.ndarray:
        """Run SIMP; returns xPhys as a float64 (nx, ny, nz) array."""
        f, fixed = self._get_bcs(force_mag)
        P = cached_problem(self, "clamp", "cpu", f, fixed, self.preserved_mask, volfrac,
                           ke=self._ke_dev)
        res = P.run(int(n_iters), volfrac=float(volfrac), tol=self.tol, max_iter=20000,
                    precision="fp32", refine=True, inner_tol=1e-3, max_outer=60,
                    warm_start=True, direct=self.solver == "direct")
        self.last_result = res
        self.last_compliance = float(res.compliance)
        return res.xPhys.cpu().numpy().reshape(self.nx, self.ny, self.nz)
