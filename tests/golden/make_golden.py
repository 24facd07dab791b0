"""Writes golden input/output vectors for the SIMP hot path by importing the
reference CPU solver (genpipeline/topology/simp_solver.py) in the development
container. Only functions that no closed form or invariant pins are dumped:

* ``SIMPSolver._get_Ke(0.3)`` -- the reference element table (it is not a valid
  hex8 stiffness, so no textbook formula reproduces it);
* ``SIMPSolver._get_bcs`` for every boundary-condition dict the reference data
  generator uses (genpipeline/topology/topo_data_gen.py:41-95), including the
  branches that fall through to defaults;
* ``SIMPSolver.run`` trajectories (SuperLU solves, OC bisection, filter) on
  small grids, one ``run(n_iters=k)`` per k.

Usage (development container only; the reference never travels):
    python tests/golden/make_golden.py /root/reference
Output: tests/golden/ref_simp_golden.npz (numpy arrays, no pickles).
"""
import io
import json
import sys
import contextlib
from pathlib import Path

import numpy as np

BCS = {
    # copied as data from genpipeline/topology/topo_data_gen.py:41-95
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
    "network": {"fixed_face": ["x_min", "y_min"], "load_face": "x_max", "load_dof": 2,
                "multi_load": True, "load_pattern": "grid"},
}
BC_GRIDS = [(8, 4, 4), (9, 5, 6)]
# (name, nx, ny, nz, bcs-key or None, volfrac, n_iters list, force, preserved?)
RUNS = [
    ("default_8x4x4", 8, 4, 4, None, 0.4, [1, 2, 3, 5], 1.0, False),
    ("lbracket_6x3x4", 6, 3, 4, "lbracket", 0.3, [1, 3], 2.0, False),
    ("preserved_8x4x4", 8, 4, 4, None, 0.35, [1, 3], 1.0, True),
]


def main(ref_root):
    sys.path.insert(0, str(ref_root))
    from genpipeline.topology.simp_solver import SIMPSolver  # noqa: E402

    out = {}
    out["ke_ref_nu0.3"] = SIMPSolver._get_Ke(0.3)
    meta = {"bcs": {}, "runs": {}}
    for (nx, ny, nz) in BC_GRIDS:
        for key, bc in BCS.items():
            s = SIMPSolver(nx=nx, ny=ny, nz=nz, boundary_conditions=dict(bc))
            f, fixed = s._get_bcs(1.5)
            tag = f"bcs/{key}/{nx}x{ny}x{nz}"
            nzi = np.nonzero(f)[0]
            out[tag + "/f_idx"] = nzi.astype(np.int64)
            out[tag + "/f_val"] = f[nzi]
            out[tag + "/fixed"] = np.asarray(fixed, dtype=np.int64)
            meta["bcs"][tag] = bc
    for name, nx, ny, nz, bkey, vf, iters, force, pres in RUNS:
        mask = None
        if pres:
            mask = np.zeros((nx, ny, nz), dtype=bool)
            mask[nx - 2:, :, :2] = True
        for k in iters:
            s = SIMPSolver(nx=nx, ny=ny, nz=nz,
                           boundary_conditions=dict(BCS[bkey]) if bkey else None,
                           preserved_mask=mask)
            with contextlib.redirect_stdout(io.StringIO()):
                xp = s.run(volfrac=vf, n_iters=k, force_mag=force)
            out[f"run/{name}/{k}/xPhys"] = xp
            out[f"run/{name}/{k}/compliance"] = np.array(s.last_compliance)
        meta["runs"][name] = {"nx": nx, "ny": ny, "nz": nz, "bcs": BCS[bkey] if bkey else None,
                              "volfrac": vf, "iters": iters, "force": force,
                              "preserved": "x>=nx-2 & z<2" if pres else None}
    dest = Path(__file__).with_name("ref_simp_golden.npz")
    np.savez_compressed(dest, **out)
    dest.with_suffix(".json").write_text(json.dumps(meta, indent=1, sort_keys=True))
    print("wrote", dest, len(out), "arrays")


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
