"""Marching cubes on the device against the CPU oracle (oracle/mesh_oracle.py),
and density_to_stl's behaviour (genpipeline/topology/mesh_export.py:7-67).

The device records are compared bit for bit: the kernel uses the oracle's
generated table (checked equal), the same vertex formula and the same
float64 normal arithmetic, rounded once to float32.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from oracle import mesh_oracle as M

pytestmark = pytest.mark.gpu


def test_mc_table_matches_oracle(gpu):
    from genpipeline_amd import _lib as L
    nt = np.zeros(256, np.int8)
    tr = np.zeros((256, 15), np.int8)
    L.call("gp_mc_tables", nt.ctypes.data_as(C.c_void_p), tr.ctypes.data_as(C.c_void_p))
    a, b = M.tables()
    assert np.array_equal(nt, a) and np.array_equal(tr, b)


@pytest.mark.parametrize("shape,level,scale", [((9, 7, 6), 0.5, (1.0, 1.0, 1.0)),
                                               ((17, 5, 12), 0.35, (100 / 16, 20 / 5, 20 / 12)),
                                               ((2, 2, 9), 0.5, (1.0, 2.0, 3.0)),
                                               ((24, 20, 18), 0.5, (1.0, 1.0, 1.0))])
def test_device_marching_cubes_matches_oracle(gpu, shape, level, scale):
    from scipy.ndimage import gaussian_filter
    from genpipeline_amd.topology.mesh_export import marching_cubes
    print(f"marching cubes {shape} level {level}", flush=True)
    rng = np.random.default_rng(sum(shape))
    d = gaussian_filter(rng.random(shape), 1.0)
    d = (d - d.min()) / (d.max() - d.min())
    out = marching_cubes(d, level, scale, device=gpu).cpu().numpy()
    ref = M.stl_records(M.marching_cubes(d, level, scale))
    assert out.shape == ref.shape and len(ref) > 0
    assert np.array_equal(out[:, 3:], ref[:, 3:])
    assert np.abs(out[:, :3] - ref[:, :3]).max() <= 1e-7


def test_density_to_stl_file_and_errors(gpu, tmp_path):
    from genpipeline_amd.topology.mesh_export import density_to_stl, read_stl
    n = 16
    x = np.arange(n) - (n - 1) / 2
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    d = np.clip(6.0 - np.sqrt(X**2 + Y**2 + Z**2), 0, 1)  # a solid ball
    p = tmp_path / "ball_mesh.stl"
    assert density_to_stl(d, str(p), voxel_size_mm=(2.0, 2.0, 2.0)) == str(p)
    rec = read_stl(str(p))
    assert np.array_equal(rec, M.stl_records(M.marching_cubes(d, 0.5, (2.0, 2.0, 2.0))))
    assert density_to_stl(np.zeros((4, 4, 4)), str(tmp_path / "e.stl")) is None
    with pytest.raises(ValueError):
        density_to_stl(np.full((4, 4, 4), 0.7), str(tmp_path / "s.stl"))
    with pytest.raises(ValueError):
        density_to_stl(np.full((4, 4, 4), 0.05), str(tmp_path / "t.stl"))
