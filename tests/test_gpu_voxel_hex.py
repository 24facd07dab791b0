"""The device voxel mesh behind the CalculiX path: decks byte-equal to the
ones the reference wrote (tests/golden/make_golden_inp.py), the solid-voxel
scan equal to np.where (the reference's scanner semantics,
cuda_kernels/__init__.py:371-388), and the first-touch node numbering equal
to the oracle's on grids of evaluator size."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import inp_oracle

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden" / "inp"


def test_decks_equal_reference_decks(gpu, tmp_path):
    from genpipeline_amd.fem import VoxelHexMesher
    cases = json.loads((GOLD / "cases.json").read_text())
    with np.load(GOLD / "fields.npz", allow_pickle=False) as d:
        fields = {k: np.array(d[k]) for k in d.files}
    for name, c in cases.items():
        print("[deck]", name, flush=True)
        out = tmp_path / f"{name}.inp"
        if c["error"] is None:
            VoxelHexMesher.voxels_to_inp(fields[name], output_path=str(out), **c["kwargs"])
            assert out.read_text() == (GOLD / f"{name}.inp").read_text(), name
        else:
            with pytest.raises(ValueError) as e:
                VoxelHexMesher.voxels_to_inp(fields[name], output_path=str(out), **c["kwargs"])
            assert str(e.value) == c["error"], name


@pytest.mark.parametrize("shape", [(33, 20, 9), (64, 64, 64), (1, 1, 300)])
def test_solid_scan_equals_where(gpu, shape):
    from genpipeline_amd.fem import get_solid_voxels
    rng = np.random.default_rng(sum(shape))
    v = rng.uniform(size=shape).astype(np.float32)
    v[v > 0.97] = 0.5  # exactly the threshold: not solid
    got = get_solid_voxels(v)
    ref = np.column_stack(np.where(v > 0.5))
    assert got.dtype == np.int64 and np.array_equal(got, ref)


@pytest.mark.parametrize("shape,frac", [((24, 16, 20), 0.45), ((32, 32, 32), 0.3)])
def test_node_numbering_equals_oracle(gpu, shape, frac):
    from genpipeline_amd.fem.voxel_hex import hex_mesh
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(7)
    g = gaussian_filter(rng.standard_normal(shape), 1.5)
    v = (g > np.quantile(g, 1 - frac)).astype(np.float32)
    elems, ijk, vox = hex_mesh(v)
    re, rc, rv = inp_oracle.mesh(v)
    assert np.array_equal(elems, np.asarray(re)) and np.array_equal(ijk, np.asarray(rc))
    D, H, W = shape
    assert np.array_equal(vox, np.asarray([(i * H + j) * W + k for i, j, k in rv]))
