"""FP64 CPU oracle for the SIMP inner loop -- TEST INFRASTRUCTURE ONLY.

This module restates, step by step, the reference CPU SIMP solver
(genpipeline/topology/simp_solver.py) and the semantics of the reference GPU
facade (genpipeline/topology/simp_solver_gpu.py) in NumPy/SciPy FP64. It is
the checker the HIP path is compared against. Only ``tests/``, the package's
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product (``genpipeline_amd``) never does.

Switches beyond the literal reference (each documented where it is used):

* ``ke``     -- ``"hex8"``: exact 2x2x2-Gauss trilinear hex stiffness (derived
  here); ``"ref"``: the reference's own 24x24 table as dumped by
  ``tests/golden/make_golden.py`` (simp_solver.py:284-1543; it is asymmetric
  and indefinite, so only the direct solve is meaningful with it).
* ``solver`` -- ``"direct"`` (SuperLU, simp_solver.py:213) or ``"pcg"``
  (FP64 Jacobi-PCG, relative residual test, zero or warm start).
* ``law``    -- ``"clamp"``: E = max(x,1e-3)^p (simp_solver.py:201-203,218);
  ``"emin"``: E = Emin + x^p (E0-Emin) (modified SIMP used by the configs).
* ``oc``     -- ``"cpu"`` constants (simp_solver.py:225-249) or ``"gpu"``
  constants (simp_solver_gpu.py:801-841).

How each function is pinned (tests/test_oracle.py):
  ke_hex8         symmetry, PSD, exactly 6 rigid-body zero modes, patch test,
                  closed-form axial bar compliance (nu=0), 45-nnz modal form.
  ke_ref          equals the golden table from the reference import.
  edof_matrix     brute force on 1x1x1 / 2x1x1 grids + node-multiplicity
                  invariant; also via the golden run trajectories.
  build_filter    closed-form 19-point weights, Hs brute force, H.1/Hs = 1.
  get_bcs_cpu     golden ``_get_bcs`` vectors for 14 BC dicts x 2 grids.
  get_bcs_gpu     golden ``SIMPSolverGPU._get_bcs`` vectors (20 BC dicts x 4
                  grids, float64 bit for bit, float32 to its rounding, and the
                  dicts the reference rejects) + closed-form invariants.
  solve           direct vs PCG agreement; axial KAT.
  sensitivity     u=0 => dc=0 (tests/test_aggressive_kernel.py:106-125);
                  sum_e E_e ce_e = f.u compliance identity.
  oc_update_*     golden run trajectories (cpu); golden ``SIMPSolverGPU.
                  _oc_update`` vectors on seeded (x, dc) (gpu); invariants.
  filter_dc       golden ``SIMPSolverGPU._build_filter`` H, Hs and
                  ``_filter_dc`` vectors.
  run             golden run trajectories (ke="ref", solver="direct").
  simp_gpu_run    golden ``SIMPSolverGPU.run`` runs with the reference's
                  ``_solve`` replaced by a direct solve (the reference method
                  is not runnable without CUDA, simp_solver_gpu.py:677), with
                  the reference table and with the exact hex8 table.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_GOLDEN = Path(__file__).resolve().parent.parent / "tests" / "golden" / "ref_simp_golden.npz"

# C3D8 local node order (simp_solver.py:255-267): (i,j,k) offsets of the 8 corners.
CORNERS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                    (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)])


# --------------------------------------------------------------------------- KE
def ke_hex8(nu: float = 0.3, E: float = 1.0) -> np.ndarray:
    """Exact unit-cube trilinear hex stiffness, 2x2x2 Gauss quadrature.

    Isotropic linear elasticity, engineering shear strains, DOF order
    3*local_node + component with the C3D8 node order above.
    """
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] += 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    g = np.array([-1.0, 1.0]) / np.sqrt(3.0)
    s = 2.0 * CORNERS - 1.0          # natural coordinates of the corners
    K = np.zeros((24, 24))
    for a, b, c in itertools.product(g, g, g):
        # dN/dxi for trilinear N_i = (1+s_x xi)(1+s_y eta)(1+s_z zeta)/8 ; dx = dxi/2
        dN = np.empty((8, 3))
        dN[:, 0] = s[:, 0] * (1 + s[:, 1] * b) * (1 + s[:, 2] * c) / 8 * 2
        dN[:, 1] = s[:, 1] * (1 + s[:, 0] * a) * (1 + s[:, 2] * c) / 8 * 2
        dN[:, 2] = s[:, 2] * (1 + s[:, 0] * a) * (1 + s[:, 1] * b) / 8 * 2
        B = np.zeros((6, 24))
        for i in range(8):
            dx, dy, dz = dN[i]
            B[0, 3 * i] = dx
            B[1, 3 * i + 1] = dy
            B[2, 3 * i + 2] = dz
            B[3, 3 * i], B[3, 3 * i + 1] = dy, dx      # gamma_xy
            B[4, 3 * i + 1], B[4, 3 * i + 2] = dz, dy  # gamma_yz
            B[5, 3 * i], B[5, 3 * i + 2] = dz, dx      # gamma_xz
        K += B.T @ D @ B * 0.125                       # detJ * weight = (1/2)^3
    return K


def ke_ref() -> np.ndarray:
    """Reference element table at nu=0.3 (simp_solver.py:284-1543), from the golden dump."""
    with np.load(_GOLDEN, allow_pickle=False) as d:
        return np.array(d["ke_ref_nu0.3"])


def get_ke(kind: str) -> np.ndarray:
    if kind == "hex8":
        return ke_hex8(0.3)
    if kind == "ref":
        return ke_ref()
    raise ValueError(kind)


# ----------------------------------------------------------------- numbering
def node_id(i, j, k, ny, nz):
    """simp_solver.py:96,277 -- n = i(ny+1)(nz+1) + j(nz+1) + k."""
    return i * (ny + 1) * (nz + 1) + j * (nz + 1) + k


def edof_matrix(nx: int, ny: int, nz: int) -> np.ndarray:
    """edof[e, 3*c+d] = 3*node(corner c of e) + d  (simp_solver.py:251-282)."""
    n_base = np.array([node_id(a, b, c, ny, nz) for a, b, c in CORNERS])
    base = (3 * n_base[:, None] + np.arange(3)[None, :]).ravel()
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    corner0 = node_id(i, j, k, ny, nz).ravel()
    return (base[None, :] + 3 * corner0[:, None]).astype(np.int64)


# ------------------------------------------------------------------- filter
def build_filter(nx: int, ny: int, nz: int, rmin: float, weights: str = "f64"):
    """H[e1,e2] = rmin - dist for dist <= rmin inside the box r=ceil(rmin)
    (simp_solver.py:1545-1567); Hs = row sums. Returns (H csr, Hs).
    weights "f32": each weight rounded once to float32, as SIMPSolverGPU's
    _build_filter fills its value arrays (simp_solver_gpu.py:915)."""
    r = int(np.ceil(rmin))
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    e1 = i * ny * nz + j * nz + k
    rows, cols, vals = [], [], []
    for di in range(-r, r + 1):
        for dj in range(-r, r + 1):
            for dk in range(-r, r + 1):
                dist = np.sqrt(di * di + dj * dj + dk * dk)
                if dist > rmin:
                    continue
                w = rmin - dist
                if weights == "f32":
                    w = float(np.float32(w))
                if w == 0.0:
                    continue  # lil_matrix does not store explicit zeros
                i2, j2, k2 = i + di, j + dj, k + dk
                ok = (i2 >= 0) & (i2 < nx) & (j2 >= 0) & (j2 < ny) & (k2 >= 0) & (k2 < nz)
                rows.append(e1[ok])
                cols.append((i2 * ny * nz + j2 * nz + k2)[ok])
                vals.append(np.full(ok.sum(), w))
    n = nx * ny * nz
    H = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    Hs = np.asarray(H.sum(axis=1)).ravel()
    return H, Hs


# ----------------------------------------------------------------------- BCs
def _grid(nx, ny, nz):
    return np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")


def face_mask_cpu(face, ix, iy, iz, nx, ny, nz):
    """simp_solver.py:159-175 (unknown names fall through to x_max)."""
    return {"x_min": ix == 0, "x_max": ix == nx, "y_min": iy == 0, "y_max": iy == ny,
            "z_min": iz == 0, "z_max": iz == nz}.get(face, ix == nx) if isinstance(face, str) else ix == nx


def get_bcs_cpu(bcs: dict, nx: int, ny: int, nz: int, force_mag: float):
    """Restates SIMPSolver._get_bcs (simp_solver.py:81-157) branch by branch."""
    ix, iy, iz = _grid(nx, ny, nz)
    nid = node_id(ix, iy, iz, ny, nz)
    n_dof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    ff = bcs.get("fixed_face", "x_min")
    faces = ff if isinstance(ff, list) else [ff]
    fixed_nodes = np.unique(np.concatenate([nid[face_mask_cpu(fc, ix, iy, iz, nx, ny, nz)] for fc in faces]))
    fixed = np.concatenate([3 * fixed_nodes, 3 * fixed_nodes + 1, 3 * fixed_nodes + 2])
    f = np.zeros(n_dof)
    lf = bcs.get("load_face", "x_max")
    ld = bcs.get("load_dof", 2)
    multi = bcs.get("multi_load", False)
    if multi and isinstance(ld, list):
        for d in ld:
            ln = nid[face_mask_cpu(lf, ix, iy, iz, nx, ny, nz)]
            v = ln[~np.isin(ln, fixed_nodes)]
            if len(v) > 0:
                sel = v[::3]
                for n in sel:
                    f[3 * n + d] -= force_mag / len(sel)
    elif lf in ("center", "center_z"):
        dist = np.sqrt((ix - nx // 2) ** 2 + (iy - ny // 2) ** 2 + (iz - nz // 2) ** 2)
        ln = nid[dist <= 2]
        v = ln[~np.isin(ln, fixed_nodes)]
        if len(v) > 0:
            sel = v[::2]
            for n in sel:
                f[3 * n + ld] -= force_mag / len(sel)
    elif lf in ("top", "x_max_offset"):
        m = (iz == nz) if lf == "top" else ((ix == nx) & (iy >= ny // 3) & (iy <= 2 * ny // 3))
        ln = nid[m]
        v = ln[~np.isin(ln, fixed_nodes)]
        if len(v) > 0:
            f[3 * v[len(v) // 2] + np.asarray(ld)] = -force_mag
    else:
        ln = nid[face_mask_cpu(lf, ix, iy, iz, nx, ny, nz)]
        v = ln[~np.isin(ln, fixed_nodes)]
        node = v[len(v) // 2] if len(v) > 0 else ln[len(ln) // 2]
        f[3 * node + np.asarray(ld)] = -force_mag
    return f, fixed


def _gpu_dof(ld):
    """The reference adds load_dof to a torch tensor of node ids
    (simp_solver_gpu.py:246,290,309): a list there is a TypeError."""
    if isinstance(ld, (list, tuple)):
        raise TypeError("tensor + list")
    return ld


def get_bcs_gpu(bcs: dict, nx: int, ny: int, nz: int, force_mag: float):
    """Restates SIMPSolverGPU._get_bcs (simp_solver_gpu.py:138-311): the plain
    branch spreads the load evenly over all non-fixed load-face nodes."""
    ix, iy, iz = _grid(nx, ny, nz)
    nid = node_id(ix, iy, iz, ny, nz)
    n_dof = 3 * (nx + 1) * (ny + 1) * (nz + 1)
    fm = {"x_min": ix == 0, "x_max": ix == nx, "y_min": iy == 0, "y_max": iy == ny,
          "z_min": iz == 0, "z_max": iz == nz}
    ff = bcs.get("fixed_face", "x_min")
    fixed_nodes = nid[fm.get(ff, ix == 0) if isinstance(ff, str) else ix == 0]
    fixed = np.concatenate([3 * fixed_nodes, 3 * fixed_nodes + 1, 3 * fixed_nodes + 2])
    f = np.zeros(n_dof)
    lf = bcs.get("load_face", "x_max")
    ld = bcs.get("load_dof", 2)
    load_nodes = nid[fm.get(lf, ix == nx) if isinstance(lf, str) else ix == nx]
    multi = bcs.get("multi_load", False)
    positions = bcs.get("load_positions", None)
    pattern = bcs.get("load_pattern", None)
    if multi and positions:
        for pos in positions:
            mp = {"top": iz == nz, "bottom": iz == 0,
                  "mid": (iz >= nz // 3) & (iz <= 2 * nz // 3),
                  "mid_y": (iy >= ny // 3) & (iy <= 2 * ny // 3),
                  "mid_z": (iz >= nz // 3) & (iz <= 2 * nz // 3)}.get(pos, iz == nz // 2)
            pn = nid[mp]
            v = pn[~np.isin(pn, fixed_nodes)]
            if len(v) > 0:
                fpn = float(force_mag) / len(positions)
                for n in v[:: max(1, len(v) // 3)]:
                    f[3 * n + _gpu_dof(ld)] -= fpn
        return f, fixed
    if multi and pattern == "grid":
        sp_ = max(1, nz // 4)
        for zv in range(0, nz + 1, sp_):
            rn = nid[iz == zv]
            v = rn[~np.isin(rn, fixed_nodes)]
            if len(v) > 0:
                fpn = float(force_mag) / ((nz // sp_) + 1)
                for n in v[:: max(1, len(v) // 2)]:
                    f[3 * n + _gpu_dof(ld)] -= fpn
        return f, fixed
    if len(load_nodes) > 0:
        v = load_nodes[~np.isin(load_nodes, fixed_nodes)]
        if len(v) == 0:
            v = load_nodes
        f[3 * v + _gpu_dof(ld)] = -float(force_mag) / len(v)
    return f, fixed


# --------------------------------------------------------------- material law
def stiffness(xPhys, law, penal, E0=1.0, Emin=1e-9):
    if law == "clamp":
        return np.maximum(xPhys, 1e-3) ** penal
    return Emin + xPhys ** penal * (E0 - Emin)


def dstiffness(xPhys, law, penal, E0=1.0, Emin=1e-9):
    if law == "clamp":
        return penal * np.maximum(xPhys, 1e-3) ** (penal - 1)
    return penal * xPhys ** (penal - 1) * (E0 - Emin)


# ------------------------------------------------------------ operator/solve
def assemble_K(KE, Ee, edof, n_dof):
    """K = sum_e Ee * KE scattered by edof (simp_solver.py:201-205).

    The reference's index vectors come from np.kron (simp_solver.py:280-281):
    entry t = 24a+b of element e carries KE[a,b] but lands at
    (edof[e,b], edof[e,a]), i.e. the CPU solver assembles KE^T. This only
    matters for the asymmetric reference table; hex8 KE is symmetric.
    """
    t = np.arange(576)
    iK = edof[:, t % 24].ravel()
    jK = edof[:, t // 24].ravel()
    sK = (KE.ravel()[None, :] * Ee[:, None]).ravel()
    return sp.coo_matrix((sK, (iK, jK)), shape=(n_dof, n_dof)).tocsr()


def ku_matfree(KE, Ee, u, edof, fixed=None):
    """y = K u without assembly (K assembled as in assemble_K, i.e. from
    KE^T per element); fixed DOFs of y are zeroed when given."""
    ue = u[edof]
    ye = (ue @ KE) * Ee[:, None]
    y = np.bincount(edof.ravel(), weights=ye.ravel(), minlength=u.size)
    if fixed is not None:
        y[fixed] = 0.0
    return y


@dataclass
class SolveInfo:
    iters: int = 0
    relres: float = 0.0
    status: str = "ok"


def pcg(K, f, fixed, tol=1e-8, maxit=20000, x0=None):
    """FP64 Jacobi-PCG on the free DOFs (u_fixed = 0). Stops when
    ||r||_2 <= tol ||f_free||_2; pAp <= 0 is reported as breakdown."""
    n = f.size
    free = np.ones(n, bool)
    free[fixed] = False
    d = K.diagonal()
    minv = np.where(free & (d > 0), 1.0 / np.where(d > 0, d, 1.0), 0.0)
    b = np.where(free, f, 0.0)
    x = np.zeros(n) if x0 is None else np.where(free, x0, 0.0)
    r = b - K @ x
    r[~free] = 0.0
    bn = np.linalg.norm(b)
    info = SolveInfo()
    if bn == 0.0:
        return np.zeros(n), info
    if np.linalg.norm(r) <= tol * bn:
        info.relres = np.linalg.norm(r) / bn
        return x, info
    z = minv * r
    p = z.copy()
    rz = r @ z
    for it in range(1, maxit + 1):
        Ap = K @ p
        Ap[~free] = 0.0
        pAp = p @ Ap
        if not pAp > 0:
            info.status = "breakdown"
            info.iters = it
            break
        a = rz / pAp
        x += a * p
        r -= a * Ap
        rn = np.linalg.norm(r)
        info.iters = it
        if rn <= tol * bn:
            break
        z = minv * r
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
    else:
        info.status = "maxit"
    info.relres = np.linalg.norm(r) / bn
    return x, info


def solve(K, f, fixed, method="direct", tol=1e-8, maxit=20000, x0=None):
    n = f.size
    if method == "direct":
        free = np.setdiff1d(np.arange(n), fixed)
        u = np.zeros(n)
        Kc = K.tocsc()
        u[free] = spla.spsolve(Kc[free, :][:, free], f[free])
        return u, SolveInfo()
    if method == "pcg":
        return pcg(K, f, fixed, tol=tol, maxit=maxit, x0=x0)
    raise ValueError(method)


# ------------------------------------------------------------ SIMP pieces
def element_energy(KE, u, edof):
    """ce_e = u_e^T KE u_e (simp_solver.py:216-217)."""
    ue = u[edof]
    return np.sum((ue @ KE) * ue, axis=1)


def sensitivity(KE, u, edof, xPhys, law, penal, E0=1.0, Emin=1e-9):
    """dc = -dE/dx * ce (simp_solver.py:218; clamp law keeps max(x,1e-3))."""
    return -dstiffness(xPhys, law, penal, E0, Emin) * element_energy(KE, u, edof)


def filter_dc(H, Hs, dc):
    """H^T (dc/Hs) (simp_solver.py:221-223)."""
    return np.asarray(H.T @ (dc / Hs)).ravel()


def oc_update_cpu(x, dc, H, Hs, volfrac, move=0.2, preserved=None):
    """simp_solver.py:225-249: l1=0, l2=1e9, <=50 bisection steps, stop test
    evaluated after computing xnew/xPhys_new at lmid."""
    l1, l2 = 0.0, 1e9
    nsteps = 0
    for _ in range(50):
        nsteps += 1
        lmid = 0.5 * (l2 + l1)
        xnew = np.clip(x * np.sqrt(np.maximum(-dc / lmid, 1e-12)), 0.0, 1.0)
        xnew = np.clip(xnew, x - move, x + move)
        if preserved is not None:
            xnew[preserved] = 1.0
        xp = np.asarray(H @ xnew).ravel() / Hs
        if preserved is not None:
            xp[preserved] = 1.0
        if (l2 - l1) / (l1 + l2 + 1e-12) < 1e-4:
            break
        if xp.mean() > volfrac:
            l1 = lmid
        else:
            l2 = lmid
    return xnew, xp, lmid, nsteps


def oc_update_gpu(x, dc, H, Hs, volfrac, move=0.2, preserved=None):
    """simp_solver_gpu.py:801-841: l2 = 2 max(max(-dc), 1), <=30 steps, 1e-5."""
    l1, l2 = 0.0, float(max(np.max(-dc), 1.0) * 2)
    nsteps = 0
    for _ in range(30):
        nsteps += 1
        lmid = 0.5 * (l2 + l1)
        xnew = np.clip(x * np.sqrt(np.maximum(-dc / lmid, 1e-15)), 0.0, 1.0)
        xnew = np.clip(xnew, x - move, x + move)
        if preserved is not None:
            xnew[preserved] = 1.0
        xp = np.asarray(H @ xnew).ravel() / Hs
        if preserved is not None:
            xp[preserved] = 1.0
        if (l2 - l1) / (l1 + l2 + 1e-15) < 1e-5:
            break
        if xp.mean() > volfrac:
            l1 = lmid
        else:
            l2 = lmid
    return xnew, xp, lmid, nsteps


# --------------------------------------------------------------- driver
@dataclass
class OracleProblem:
    """A fully specified SIMP problem: grid, material, BCs as arrays."""
    nx: int
    ny: int
    nz: int
    f: np.ndarray
    fixed: np.ndarray
    penal: float = 3.0
    rmin: float = 1.5
    volfrac: float = 0.3
    move: float = 0.2
    E0: float = 1.0
    Emin: float = 1e-9
    law: str = "emin"
    ke: str = "hex8"
    oc: str = "cpu"
    preserved: np.ndarray | None = None
    # "cpu": element blocks are KE^T (the kron index vectors of
    # simp_solver.py:280-281); "gpu": KE itself (repeat_interleave / repeat
    # index vectors, simp_solver_gpu.py:66-67). Only an asymmetric table
    # tells them apart.
    assembly: str = "cpu"
    filter_weights: str = "f64"   # "f32": SIMPSolverGPU's filter (build_filter)
    _cache: dict = field(default_factory=dict, repr=False)

    @property
    def n_elem(self):
        return self.nx * self.ny * self.nz

    @property
    def n_dof(self):
        return 3 * (self.nx + 1) * (self.ny + 1) * (self.nz + 1)

    def setup(self):
        if not self._cache:
            self._cache["KE"] = get_ke(self.ke)
            self._cache["edof"] = edof_matrix(self.nx, self.ny, self.nz)
            self._cache["H"], self._cache["Hs"] = build_filter(self.nx, self.ny, self.nz, self.rmin,
                                                               self.filter_weights)
        c = self._cache
        return c["KE"], c["edof"], c["H"], c["Hs"]

    def solve_u(self, xPhys, method="pcg", tol=1e-8, maxit=20000, x0=None):
        KE, edof, _, _ = self.setup()
        Ee = stiffness(xPhys, self.law, self.penal, self.E0, self.Emin)
        K = assemble_K(KE if self.assembly == "cpu" else KE.T, Ee, edof, self.n_dof)
        return solve(K, self.f, self.fixed, method, tol, maxit, x0)

    def step(self, x, xPhys, method="pcg", tol=1e-8, x0=None):
        """One OC iteration from state (x, xPhys). Returns
        (xnew, xPhys_new, u, compliance_of_input, info dict)."""
        KE, edof, H, Hs = self.setup()
        u, sinfo = self.solve_u(xPhys, method, tol, x0=x0)
        dc = sensitivity(KE, u, edof, xPhys, self.law, self.penal, self.E0, self.Emin)
        dcf = filter_dc(H, Hs, dc)
        oc = oc_update_cpu if self.oc == "cpu" else oc_update_gpu
        xnew, xp, lmid, nb = oc(x, dcf, H, Hs, self.volfrac, self.move, self.preserved)
        c = float(self.f @ u)
        return xnew, xp, u, c, {"solve": sinfo, "lmid": lmid, "n_bisect": nb, "dc": dcf}

    def run(self, n_iters, method="pcg", tol=1e-8, x0_density=None, warm=False, record=False):
        """Reference driver (simp_solver.py:56-79): x0 = volfrac (preserved=1),
        xPhys0 = x0 unfiltered; n_iters x {solve, sens, filter, OC}; then a
        final solve for compliance."""
        x = np.full(self.n_elem, float(self.volfrac)) if x0_density is None else np.array(x0_density, float)
        if self.preserved is not None:
            x[self.preserved] = 1.0
        xPhys = x.copy()
        hist = []
        u = None
        for _ in range(n_iters):
            x, xPhys, u, c, info = self.step(x, xPhys, method, tol, x0=u if warm else None)
            if record:
                hist.append((x.copy(), xPhys.copy(), c, info["solve"].iters, info["n_bisect"]))
        u, _ = self.solve_u(xPhys, method, tol, x0=u if warm else None)
        comp = float(self.f @ u)
        return xPhys, comp, hist


def reference_problem(nx, ny, nz, bcs=None, force_mag=1.0, penal=3.0, rmin=1.5, volfrac=0.4,
                      preserved=None, ke="ref", law="clamp", oc="cpu", gpu_bcs=False):
    """Problem with the reference SIMPSolver defaults (simp_solver.py:24-54);
    gpu_bcs: SIMPSolverGPU's loads and assembly orientation."""
    b = {"fixed_face": "x_min", "load_face": "x_max", "load_dof": 2}
    if bcs:
        b.update(bcs)
    f, fixed = (get_bcs_gpu if gpu_bcs else get_bcs_cpu)(b, nx, ny, nz, force_mag)
    pm = None if preserved is None else np.asarray(preserved).ravel().astype(bool)
    return OracleProblem(nx, ny, nz, f, fixed, penal=penal, rmin=rmin, volfrac=volfrac,
                         law=law, ke=ke, oc=oc, preserved=pm,
                         assembly="gpu" if gpu_bcs else "cpu",
                         filter_weights="f32" if gpu_bcs else "f64")


def gpu_tol_schedule(i):
    """SIMPSolverGPU.run's per-iteration PCG tolerance (simp_solver_gpu.py:97-103)."""
    return 1e-4 if i < 3 else (1e-5 if i < 8 else 1e-6)


def simp_gpu_run(P, n_iters, volfrac, x_init=None, method="pcg", tol=None):
    """SIMPSolverGPU.run (simp_solver_gpu.py:71-136) on the oracle.

    x_init is rescaled to volfrac and clipped to [1e-3, 1] (:79-84);
    preserved elements start at 1 and xPhys0 = x0 (:86-88). Each iteration:
    a solve warm-started from the previous compliance solve (``_u_prev``,
    :112 and :747-751 with :585-705), sensitivity, NaN abort (:106-108),
    filter adjoint, GPU-constant OC, a compliance solve from zero, and the
    early stop after three consecutive relative changes below 1% (:114-125).
    ``tol`` None follows the tolerance schedule (``method`` "pcg"); a number
    fixes every solve's tolerance. Returns (xPhys, last_compliance,
    iterations done, compliance after each iteration)."""
    KE, edof, H, Hs = P.setup()
    if x_init is not None:
        x = np.asarray(x_init, np.float64).ravel()
        x = np.clip(x * (volfrac / max(x.mean(), 1e-6)), 1e-3, 1.0)
    else:
        x = np.full(P.n_elem, float(volfrac))
    if P.preserved is not None:
        x[P.preserved] = 1.0
    xPhys = x.copy()
    u_prev, prev, changes, comps, curr, done = None, np.inf, [], [], None, 0
    for i in range(n_iters):
        t = gpu_tol_schedule(i) if tol is None else tol
        u, _ = P.solve_u(xPhys, method, t, x0=u_prev)
        dc = sensitivity(KE, u, edof, xPhys, P.law, P.penal, P.E0, P.Emin)
        if np.isnan(dc).any():
            break
        x, xPhys, _, _ = oc_update_gpu(x, filter_dc(H, Hs, dc), H, Hs, volfrac, P.move,
                                       P.preserved)
        done = i + 1
        uc, _ = P.solve_u(xPhys, method, 1e-6 if tol is None else tol)
        u_prev = uc
        curr = float(P.f @ uc)
        comps.append(curr)
        if np.isfinite(prev):
            changes.append(abs(prev - curr) / (prev + 1e-10))
        if len(changes) >= 3 and all(c < 0.01 for c in changes[-3:]):
            break
        prev = curr
    return xPhys, curr, done, comps
