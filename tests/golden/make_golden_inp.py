"""Writes golden CalculiX decks of the reference's VoxelHexMesher
(genpipeline/fem/voxel_fem.py:74-226) by importing it in the development
container. The reference's tests pin only node counts and card names
(tests/test_voxel_fem.py:84-135); the node numbering, number formats and
card order of a whole deck are pinned by these files.

The reference meshes the solid voxels found by get_solid_voxels_simd
(cuda_kernels/__init__.py:371-388), which JIT-compiles an AVX-512 scanner
and otherwise uses np.where(voxels > 0.5). This script replaces that one
function by the reference's own np.where branch (:385-388), which lists the
same voxels in the same C order, so nothing is compiled.

Usage (development container only; the reference never travels):
    python tests/golden/make_golden_inp.py /root/reference
Output: tests/golden/inp/<case>.inp, the fields in
tests/golden/inp/fields.npz and the case parameters in cases.json; error
cases are recorded with the exception's message.
"""
import json
import sys
from pathlib import Path

import numpy as np
from scipy.ndimage import gaussian_filter

OUT = Path(__file__).resolve().parent / "inp"


def fields():
    rng = np.random.default_rng(4)
    out = {}
    out["cube2"] = (np.ones((2, 2, 2), np.float32), {})
    out["single"] = (np.ones((1, 1, 1), np.float32), {})
    f = (rng.uniform(size=(6, 5, 4)) > 0.35).astype(np.float32)
    f[0, :, :] = 1.0
    out["holes_6x5x4_bbox"] = (f, {"bbox": {"x": [0.0, 30.0], "y": [-5.0, 5.0], "z": [0.0, 8.0]},
                                   "fixed_face": "y_min", "load_face": "z_max",
                                   "force_n": 500.0})
    g = gaussian_filter(rng.standard_normal((7, 6, 5)), 1.0)
    out["smooth_7x6x5"] = ((g > np.quantile(g, 0.55)).astype(np.float32) * 0.9 + 0.05, {})
    h = rng.uniform(size=(5, 4, 6)).astype(np.float32)
    out["graded_5x4x6_xmax"] = (h, {"fixed_face": "x_max", "load_face": "x_min",
                                    "E_mpa": 70000.0, "poisson": 0.33})
    out["void"] = (np.zeros((3, 3, 3), np.float32), {})
    inner = np.zeros((4, 4, 4), np.float32)
    inner[1:3, 1:3, 1:3] = 1.0
    out["no_face_nodes"] = (inner, {})
    return out


def main(ref_root):
    sys.path.insert(0, str(ref_root))
    import genpipeline.cuda_kernels as ck
    from genpipeline.fem.voxel_fem import VoxelHexMesher

    def solid_where(voxels):  # the reference's own fallback branch, :385-388
        return np.column_stack(np.where(voxels > 0.5)).astype(np.int32)

    ck.get_solid_voxels_simd = solid_where
    OUT.mkdir(exist_ok=True)
    cases, arrays = {}, {}
    for name, (vox, kw) in fields().items():
        arrays[name] = vox
        path = OUT / f"{name}.inp"
        try:
            VoxelHexMesher.voxels_to_inp(vox, output_path=str(path), **kw)
            cases[name] = {"kwargs": kw, "error": None}
        except ValueError as e:
            cases[name] = {"kwargs": kw, "error": str(e)}
    np.savez_compressed(OUT / "fields.npz", **arrays)
    (OUT / "cases.json").write_text(json.dumps(cases, indent=1, sort_keys=True))
    print("wrote", len(cases), "cases to", OUT)


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
