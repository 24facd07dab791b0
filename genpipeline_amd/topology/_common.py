"""Shared host helpers of the solver facades."""
from __future__ import annotations

import hashlib

import numpy as np

from ..problem import Problem, ProblemSpec

# C3D8 corner offsets (simp_solver.py:255-267)
_CORNERS = ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1))


def build_edof_mapping(nx: int, ny: int, nz: int):
    """edof (n_elem x 24, int64) and the Kronecker-expanded COO index vectors
    iK = kron(edof, ones(24,1)), jK = kron(edof, ones(1,24)) in the layout of
    simp_solver.py:251-282. Host-side only, for callers that read them."""
    sI, sJ = (ny + 1) * (nz + 1), nz + 1
    base = np.array([3 * (a * sI + b * sJ + c) + d for (a, b, c) in _CORNERS for d in range(3)],
                    dtype=np.int64)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    n0 = (i * sI + j * sJ + k).ravel().astype(np.int64)
    edof = base[None, :] + 3 * n0[:, None]
    # entry t = 24 a + b of element e: iK = edof[e, b], jK = edof[e, a]
    iK = np.tile(edof, (1, 24)).ravel()
    jK = np.repeat(edof, 24, axis=1).ravel()
    return edof, iK, jK


def _digest(*arrays) -> str:
    h = hashlib.sha1()
    for a in arrays:
        if a is None:
            h.update(b"-")
        else:
            a = np.ascontiguousarray(a)
            h.update(str(a.dtype).encode())
            h.update(a.tobytes())
    return h.hexdigest()


def cached_problem(owner, law, oc, f, fixed, preserved, volfrac, penal=None, E0=1.0, Emin=1e-9,
                   filter_weights="f64", ke=None):
    """The owner's device problem for these BCs, created on first use and
    reused while grid, material, element table, loads and supports stay the
    same. ke: None (exact hex8) or a 24x24 table in the device orientation
    (y_e = ke u_e)."""
    penal = owner.penal if penal is None else penal
    if ke is not None:
        ke = np.ascontiguousarray(ke, np.float64).reshape(24, 24)
    key = (law, oc, filter_weights, float(penal), float(owner.rmin), float(E0), float(Emin),
           None if ke is None else ke.tobytes(),
           _digest(np.asarray(f, np.float64), np.sort(np.asarray(fixed, np.int64)),
                   None if preserved is None else np.asarray(preserved, bool)))
    P = getattr(owner, "_problem", None)
    if P is not None and getattr(owner, "_problem_key", None) == key:
        return P
    if P is not None:
        P.close()
    spec = ProblemSpec(owner.nx, owner.ny, owner.nz, penal=float(penal), rmin=float(owner.rmin),
                       volfrac=float(volfrac), E0=float(E0), Emin=float(Emin), law=law, oc=oc,
                       filter_weights=filter_weights)
    P = Problem(spec, fixed, f, preserved=preserved, ke=ke,
                device=getattr(owner, "_torch_device", None))
    owner._problem = P
    owner._problem_key = key
    return P
