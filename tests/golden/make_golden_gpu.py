"""Writes golden input/output vectors for the reference GPU SIMP solver
(genpipeline/topology/simp_solver_gpu.py) by importing it on CPU torch in the
development container. These are the functions no closed form pins:

* ``SIMPSolverGPU._get_bcs`` (:138-311) for every boundary-condition dict of
  the reference data generator (topo_data_gen.py:39-94), including the
  ``load_positions`` (branch, y_tree) and ``grid`` (network) patterns, in
  float64 and float32; dicts the reference rejects are recorded with the
  exception's type name;
* ``SIMPSolverGPU._build_filter`` (:869-930) H and Hs, and ``_filter_dc``
  (:796-799) on seeded sensitivities;
* ``SIMPSolverGPU._oc_update`` (:801-841) on seeded (x, dc), with and
  without a preserved mask;
* ``SIMPSolverGPU.run`` (:71-136) end to end. Without CUDA the reference's
  ``_solve`` reaches a branch that reads an undefined name (:677), so each run
  replaces that one method with a SciPy direct solve of the same system
  (u = 0 on the fixed DOFs, ``_u_prev`` kept as every reference branch keeps
  it). The runs use the reference's own element table, and also -- as data,
  the attribute ``Ke`` set before ``run`` -- the exact hex8 matrix the build
  uses, so the GPU facade's driver can be compared with them.

Usage (development container only; the reference never travels):
    python tests/golden/make_golden_gpu.py /root/reference
Output: tests/golden/ref_simp_gpu_golden.npz (numpy arrays, no pickles) and
ref_simp_gpu_golden.json (the case parameters).
"""
import contextlib
import io
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.simp_oracle import ke_hex8  # noqa: E402

BCS = {
    # the reference data generator's table (topo_data_gen.py:39-94), as data
    "cantilever": {"fixed_face": "x_min", "load_face": "x_max", "load_dof": 2},
    "lbracket": {"fixed_face": "z_min", "load_face": "x_max", "load_dof": 2},
    "bridge": {"fixed_face": ["x_min", "x_max"], "load_face": "center", "load_dof": 2},
    "arch": {"fixed_face": ["x_min", "x_max"], "load_face": "top", "load_dof": 2},
    "3dload": {"fixed_face": "x_min", "load_face": "x_max", "load_dof": [1, 2], "multi_load": True},
    "offset": {"fixed_face": "x_min", "load_face": "x_max_offset", "load_dof": 2},
    "center": {"fixed_face": "x_min", "load_face": "center_z", "load_dof": 2},
    "portal": {"fixed_face": ["y_min", "y_max"], "load_face": "top_center", "load_dof": 2},
    "spoke": {"fixed_face": "center", "load_face": "outer_ring", "load_dof": 2},
    "tower": {"fixed_face": "z_min", "load_face": "z_max", "load_dof": [1, 2]},
    "sandwich": {"fixed_face": ["z_min", "z_max"], "load_face": "center", "load_dof": 2},
    "simply": {"fixed_face": ["x_min", "x_max"], "load_face": "top", "load_dof": 2},
    "branch": {"fixed_face": "x_min", "load_face": "x_max", "load_dof": 2, "multi_load": True,
               "load_positions": ["top", "mid", "bottom"]},
    "y_tree": {"fixed_face": "z_min", "load_face": "x_max", "load_dof": 2, "multi_load": True,
               "load_positions": ["top", "mid_y", "mid_z"]},
    "network": {"fixed_face": ["x_min", "y_min"], "load_face": "x_max", "load_dof": 2,
                "multi_load": True, "load_pattern": "grid"},
    # single-face variants of the other branches' faces
    "yface": {"fixed_face": "y_max", "load_face": "y_min", "load_dof": 0},
    "zface": {"fixed_face": "z_max", "load_face": "z_min", "load_dof": 1},
    "xmaxfix": {"fixed_face": "x_max", "load_face": "x_min", "load_dof": 2},
    "mixedpos": {"fixed_face": "x_min", "load_face": "x_max", "load_dof": 1, "multi_load": True,
                 "load_positions": ["bottom", "other"]},
    "listpos": {"fixed_face": "x_min", "load_face": "x_max", "load_dof": [0, 2], "multi_load": True,
                "load_positions": ["top"]},
}
BC_GRIDS = [(8, 4, 4), (9, 5, 6), (12, 6, 6), (5, 13, 3)]
BC_FORCE = 1.5
# (nx, ny, nz, rmin) for the filter and OC vectors
FILTER_CASES = [(8, 4, 4, 1.5), (9, 5, 6, 1.5), (7, 5, 6, 2.5), (6, 6, 6, 2.0)]
# (name, nx, ny, nz, bcs key, volfrac, n_iters, force, x_init seed or None,
#  preserved, element table)
RUNS = [
    ("default_10x5x4", 10, 5, 4, None, 0.35, 40, 1.0, None, False, "hex8"),
    ("xinit_10x5x4", 10, 5, 4, None, 0.35, 12, 1.0, 5, False, "hex8"),
    ("preserved_8x4x4", 8, 4, 4, None, 0.35, 10, 1.0, None, True, "hex8"),
    ("branch_9x5x6", 9, 5, 6, "branch", 0.3, 8, 1.0, None, False, "hex8"),
    ("network_9x5x6", 9, 5, 6, "network", 0.3, 8, 2.0, None, False, "hex8"),
    ("lbracket_8x4x6", 8, 4, 6, "lbracket", 0.3, 6, 1.0, 11, False, "hex8"),
    ("default_8x4x4_reftable", 8, 4, 4, None, 0.4, 5, 1.0, None, False, "ref"),
]


def _direct_solve(self, K, f, fixed_dofs, tol=1e-6, max_iter=2000, x0=None):
    """Stand-in for SIMPSolverGPU._solve on CPU: SciPy direct solve of the
    free block of the reference's own assembled K (u = 0 on fixed DOFs)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    import torch
    n = K.shape[0]
    Kc = sp.csr_matrix((K.values().numpy(), K.col_indices().numpy(), K.crow_indices().numpy()),
                       shape=(n, n))
    fd = np.unique(fixed_dofs.numpy())
    free = np.setdiff1d(np.arange(n), fd)
    u = np.zeros(n)
    Kf = Kc[free, :][:, free].tocsc()
    u[free] = spla.spsolve(Kf, f.numpy().astype(np.float64)[free])
    x = torch.from_numpy(u).to(K.dtype)
    self._u_prev = x.clone()
    return x


def main(ref_root):
    sys.path.insert(0, str(ref_root))
    import torch
    from genpipeline.topology.simp_solver_gpu import SIMPSolverGPU  # noqa: E402

    out, meta = {}, {"bcs": {}, "bcs_errors": {}, "filter": {}, "oc": {}, "runs": {}}
    # ------------------------------------------------------------ BCs
    for (nx, ny, nz) in BC_GRIDS:
        for key, bc in BCS.items():
            for dt, dname in ((torch.float64, "f64"), (torch.float32, "f32")):
                tag = f"bcs/{key}/{nx}x{ny}x{nz}/{dname}"
                s = SIMPSolverGPU(nx=nx, ny=ny, nz=nz, boundary_conditions=dict(bc), device="cpu",
                                  dtype=dt)
                try:
                    f, fixed = s._get_bcs(BC_FORCE)
                except Exception as e:  # the reference rejects this dict on this grid
                    meta["bcs_errors"][tag] = type(e).__name__
                    continue
                f = f.numpy()
                nzi = np.nonzero(f)[0]
                out[tag + "/f_idx"] = nzi.astype(np.int64)
                out[tag + "/f_val"] = f[nzi]
                out[tag + "/fixed"] = fixed.numpy().astype(np.int64)
                meta["bcs"][tag] = bc
    # ----------------------------------------------------- filter + OC
    for ci, (nx, ny, nz, rmin) in enumerate(FILTER_CASES):
        n = nx * ny * nz
        tag = f"filter/{nx}x{ny}x{nz}_r{rmin}"
        s = SIMPSolverGPU(nx=nx, ny=ny, nz=nz, rmin=rmin, device="cpu", dtype=torch.float64)
        H = s._H.coalesce()
        out[tag + "/H_idx"] = H.indices().numpy().astype(np.int64)
        out[tag + "/H_val"] = H.values().numpy()
        out[tag + "/Hs"] = s._Hs.numpy()
        rng = np.random.default_rng(100 + ci)
        dc = -rng.uniform(0.0, 4.0, n)
        out[tag + "/dc_in"] = dc
        out[tag + "/dc_out"] = s._filter_dc(torch.from_numpy(dc)).numpy()
        H32, Hs32 = s._build_filter(nx, ny, nz, rmin, torch.float32)
        out[tag + "/Hs_f32"] = Hs32.numpy()
        meta["filter"][tag] = {"nx": nx, "ny": ny, "nz": nz, "rmin": rmin}
        for pres in (False, True):
            for vf in (0.3, 0.5):
                otag = f"oc/{nx}x{ny}x{nz}_r{rmin}/vf{vf}/{'pres' if pres else 'nopres'}"
                mask = None
                if pres:
                    mask = np.zeros((nx, ny, nz), bool)
                    mask[nx - 2:, :, :2] = True
                so = SIMPSolverGPU(nx=nx, ny=ny, nz=nz, rmin=rmin, device="cpu",
                                   dtype=torch.float64, preserved_mask=mask)
                x = rng.uniform(0.05, 0.95, n)
                if mask is not None:
                    x[mask.ravel()] = 1.0
                dco = -rng.uniform(0.0, 5.0, n) * rng.uniform(0.0, 1.0, n) ** 3
                xt = torch.from_numpy(x)
                xnew, xphys = so._oc_update(xt, xt.clone(), torch.from_numpy(dco), vf)
                out[otag + "/x"] = x
                out[otag + "/dc"] = dco
                out[otag + "/xnew"] = xnew.numpy()
                out[otag + "/xphys"] = xphys.numpy()
                meta["oc"][otag] = {"nx": nx, "ny": ny, "nz": nz, "rmin": rmin, "volfrac": vf,
                                    "preserved": "x>=nx-2 & z<2" if pres else None}
    # ------------------------------------------------------------- runs
    for name, nx, ny, nz, bkey, vf, n_iters, force, xseed, pres, table in RUNS:
        mask = None
        if pres:
            mask = np.zeros((nx, ny, nz), bool)
            mask[nx - 2:, :, :2] = True
        SIMPSolverGPU._RESOURCE_CACHE.clear()
        s = SIMPSolverGPU(nx=nx, ny=ny, nz=nz, boundary_conditions=dict(BCS[bkey]) if bkey else None,
                          device="cpu", dtype=torch.float64, preserved_mask=mask)
        s._solve = _direct_solve.__get__(s)
        if table == "hex8":
            s.Ke = torch.from_numpy(ke_hex8(0.3))
        comps = []
        calc = s._calculate_compliance

        def recorded(xPhys, fm, _calc=calc, _c=comps):
            c = _calc(xPhys, fm)
            _c.append(c)
            return c

        s._calculate_compliance = recorded
        x_init = None
        if xseed is not None:
            x_init = np.random.default_rng(xseed).uniform(0.05, 1.0, (nx, ny, nz))
            out[f"run/{name}/x_init"] = x_init
        with contextlib.redirect_stdout(io.StringIO()):
            xp = s.run(volfrac=vf, n_iters=n_iters, force_mag=force, x_init=x_init)
        out[f"run/{name}/xPhys"] = xp
        out[f"run/{name}/compliance"] = np.array(s.last_compliance)
        out[f"run/{name}/compliance_history"] = np.array(comps)
        meta["runs"][name] = {"nx": nx, "ny": ny, "nz": nz, "bcs": BCS[bkey] if bkey else None,
                              "volfrac": vf, "n_iters": n_iters, "force": force,
                              "x_init_seed": xseed, "iters_done": len(comps),
                              "preserved": "x>=nx-2 & z<2" if pres else None, "table": table}
    dest = Path(__file__).with_name("ref_simp_gpu_golden.npz")
    np.savez_compressed(dest, **out)
    dest.with_suffix(".json").write_text(json.dumps(meta, indent=1, sort_keys=True))
    print("wrote", dest, len(out), "arrays;", len(meta["bcs_errors"]), "rejected BC cases")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
