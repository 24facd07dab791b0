"""Boundary conditions of the reference solvers, as host-side index arithmetic.

``simp_bcs`` follows ``SIMPSolver._get_bcs`` (genpipeline/topology/
simp_solver.py:81-175) and ``simp_gpu_bcs`` follows ``SIMPSolverGPU._get_bcs``
(genpipeline/topology/simp_solver_gpu.py:138-311). Both return
``(f, fixed_dofs)`` with f a float64 load vector of length n_dof and
fixed_dofs an int64 array in the reference's order. The two differ on
purpose (the reference CPU solver loads one node, the GPU solver spreads the
load over the whole face); each facade uses its own.
"""
from __future__ import annotations

import numpy as np


def _nodes(nx, ny, nz):
    """(i, j, k, node_id) over the node grid in C order (meshgrid 'ij' ravel)."""
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    nid = i * (ny + 1) * (nz + 1) + j * (nz + 1) + k
    return i, j, k, nid


def _face(name, i, j, k, nx, ny, nz, default):
    table = {"x_min": i == 0, "x_max": i == nx, "y_min": j == 0, "y_max": j == ny,
             "z_min": k == 0, "z_max": k == nz}
    if isinstance(name, str) and name in table:
        return table[name]
    return default


def _dofs_of(nodes):
    nodes = np.asarray(nodes, dtype=np.int64)
    return np.concatenate([3 * nodes, 3 * nodes + 1, 3 * nodes + 2])


def simp_bcs(bcs: dict, nx: int, ny: int, nz: int, force_mag: float):
    """SIMPSolver._get_bcs: fixed = all DOFs of the fixed face(s) (np.unique
    over a list of faces; unknown names act as x_max); load branches
    multi_load / center / center_z / top / x_max_offset / plain face."""
    i, j, k, nid = _nodes(nx, ny, nz)
    x_max = i == nx
    ff = bcs.get("fixed_face", "x_min")
    faces = ff if isinstance(ff, list) else [ff]
    fixed_nodes = np.unique(np.concatenate([nid[_face(fc, i, j, k, nx, ny, nz, x_max)] for fc in faces]))
    fixed = _dofs_of(fixed_nodes)
    f = np.zeros(3 * nid.size)
    lf = bcs.get("load_face", "x_max")
    ld = bcs.get("load_dof", 2)

    def free(mask):
        cand = nid[mask]
        return cand, cand[~np.isin(cand, fixed_nodes)]

    if bcs.get("multi_load", False) and isinstance(ld, list):
        for d in ld:
            _, ok = free(_face(lf, i, j, k, nx, ny, nz, x_max))
            pick = ok[::3]
            if ok.size:
                for n in pick:  # sequential -= keeps the reference's rounding
                    f[3 * n + d] -= force_mag / pick.size
    elif lf in ("center", "center_z"):
        r = np.sqrt((i - nx // 2) ** 2 + (j - ny // 2) ** 2 + (k - nz // 2) ** 2)
        _, ok = free(r <= 2)
        pick = ok[::2]
        if ok.size:
            for n in pick:
                f[3 * n + ld] -= force_mag / pick.size
    elif lf == "top" or lf == "x_max_offset":
        mask = (k == nz) if lf == "top" else (x_max & (j >= ny // 3) & (j <= 2 * ny // 3))
        _, ok = free(mask)
        if ok.size:
            f[3 * ok[ok.size // 2] + np.asarray(ld)] = -force_mag
    else:
        cand, ok = free(_face(lf, i, j, k, nx, ny, nz, x_max))
        node = ok[ok.size // 2] if ok.size else cand[cand.size // 2]
        f[3 * node + np.asarray(ld)] = -force_mag
    return f, fixed


def _scalar_dof(ld):
    """SIMPSolverGPU adds load_dof to a tensor of node ids; a list there
    raises TypeError (simp_solver_gpu.py:246, 290, 309), so a list-valued
    load_dof is rejected wherever the reference would apply it."""
    if isinstance(ld, (list, tuple)):
        raise TypeError("SIMPSolverGPU._get_bcs takes a single load_dof, got a list")
    return int(ld)


def simp_gpu_bcs(bcs: dict, nx: int, ny: int, nz: int, force_mag: float):
    """SIMPSolverGPU._get_bcs: one fixed face (unknown -> x_min), load spread
    evenly over the non-fixed nodes of the load face (unknown -> x_max), plus
    the multi_load 'load_positions' and 'grid' patterns. A list load_dof
    raises TypeError as in the reference."""
    i, j, k, nid = _nodes(nx, ny, nz)
    ff = bcs.get("fixed_face", "x_min")
    fixed_nodes = nid[_face(ff, i, j, k, nx, ny, nz, i == 0)]
    fixed = _dofs_of(fixed_nodes)
    f = np.zeros(3 * nid.size)
    lf = bcs.get("load_face", "x_max")
    ld = bcs.get("load_dof", 2)
    load_nodes = nid[_face(lf, i, j, k, nx, ny, nz, i == nx)]
    multi = bcs.get("multi_load", False)
    positions = bcs.get("load_positions", None)
    if multi and positions:
        for pos in positions:
            masks = {"top": k == nz, "bottom": k == 0, "mid": (k >= nz // 3) & (k <= 2 * nz // 3),
                     "mid_y": (j >= ny // 3) & (j <= 2 * ny // 3),
                     "mid_z": (k >= nz // 3) & (k <= 2 * nz // 3)}
            cand = nid[masks.get(pos, k == nz // 2)]
            ok = cand[~np.isin(cand, fixed_nodes)]
            if ok.size:
                share = float(force_mag) / len(positions)
                for n in ok[:: max(1, ok.size // 3)]:
                    f[3 * n + _scalar_dof(ld)] -= share
        return f, fixed
    if multi and bcs.get("load_pattern", None) == "grid":
        step = max(1, nz // 4)
        for zv in range(0, nz + 1, step):
            cand = nid[k == zv]
            ok = cand[~np.isin(cand, fixed_nodes)]
            if ok.size:
                share = float(force_mag) / ((nz // step) + 1)
                for n in ok[:: max(1, ok.size // 2)]:
                    f[3 * n + _scalar_dof(ld)] -= share
        return f, fixed
    if load_nodes.size:
        ok = load_nodes[~np.isin(load_nodes, fixed_nodes)]
        if ok.size == 0:
            ok = load_nodes
        f[3 * ok + _scalar_dof(ld)] = -float(force_mag) / ok.size
    return f, fixed


def clamp_x0(nx: int, ny: int, nz: int) -> np.ndarray:
    """All DOFs of the nodes on face i=0 (node order), 3(ny+1)(nz+1) DOFs."""
    n = np.arange((ny + 1) * (nz + 1), dtype=np.int64)  # i=0 nodes are 0..(ny+1)(nz+1)-1
    return _dofs_of(n)
