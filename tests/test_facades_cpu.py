"""Host-side behaviour of the drop-in facades (no GPU): signatures and
defaults of the reference interfaces, index layouts, boundary conditions,
the evaluator's penalty path and the pipeline utilities."""
import inspect
import json

import numpy as np
import pytest

from oracle import simp_oracle as O


def _defaults(fn):
    return {k: v.default for k, v in inspect.signature(fn).parameters.items()
            if v.default is not inspect.Parameter.empty and v.kind != v.KEYWORD_ONLY}


def test_reference_signatures_and_defaults():
    import torch
    from genpipeline_amd.fem import GPUConjugateGradientFEM, GPUVoxelFEM, VoxelFEMEvaluator
    from genpipeline_amd.topology import DenseGPUSolver, SIMPSolver, SIMPSolverGPU

    # simp_solver.py:24-33, :55-57
    assert _defaults(SIMPSolver.__init__) == dict(nx=32, ny=8, nz=8, penal=3.0, rmin=1.5,
                                                   boundary_conditions=None, preserved_mask=None)
    assert _defaults(SIMPSolver.run) == dict(volfrac=0.4, n_iters=80, force_mag=1.0)
    # simp_solver_gpu.py:16-27, :71-77
    assert _defaults(SIMPSolverGPU.__init__) == dict(
        nx=32, ny=8, nz=8, penal=3.0, rmin=1.5, boundary_conditions=None, device="cuda",
        preserved_mask=None, dtype=torch.float32)
    assert _defaults(SIMPSolverGPU.run) == dict(volfrac=0.4, n_iters=15, force_mag=1.0, x_init=None)
    assert _defaults(SIMPSolverGPU._solve) == dict(tol=1e-6, max_iter=2000, x0=None)
    # simp_solver_dense_gpu.py:23-31, :238-240
    assert _defaults(DenseGPUSolver.__init__) == dict(nx=32, ny=8, nz=8, penal=3.0, rmin=1.5,
                                                       device="cuda")
    assert _defaults(DenseGPUSolver.run) == dict(volfrac=0.25, n_iters=15, force_mag=1000)
    # gpu_fem_solver.py:43-51, alias :137
    assert _defaults(GPUVoxelFEM.__init__) == dict(fixed_face="x_min", load_face="x_max",
                                                    load_dof=2, force_n=1000.0, device="cuda")
    assert GPUConjugateGradientFEM is GPUVoxelFEM
    # voxel_fem.py:406-415
    assert _defaults(VoxelFEMEvaluator.__init__) == dict(
        ccx_cmd=None, output_dir=None, fixed_face="x_min", load_face="x_max", force_n=1000.0,
        vae_model=None, use_gpu=False)
    assert _defaults(VoxelFEMEvaluator.evaluate) == dict(bbox=None)


def test_simp_solver_host_state_matches_reference_layout(tmp_path):
    from genpipeline_amd.topology import SIMPSolver
    s = SIMPSolver(nx=3, ny=2, nz=2, boundary_conditions={"load_dof": 1},
                   preserved_mask=np.ones((3, 2, 2)))
    assert s.bcs == {"fixed_face": "x_min", "load_face": "x_max", "load_dof": 1}
    assert s.preserved_mask.shape == (12,) and s.preserved_mask.all()
    assert np.abs(s.Ke - O.ke_hex8(0.3)).max() < 1e-15
    assert np.abs(SIMPSolver._get_Ke(0.25) - O.ke_hex8(0.25)).max() < 1e-15
    edof, iK, jK = s._build_edof_mapping()
    ed = O.edof_matrix(3, 2, 2)
    assert np.array_equal(edof, ed)
    # np.kron(edof, ones((24,1))) / np.kron(edof, ones((1,24))) (simp_solver.py:280-281)
    assert np.array_equal(iK, np.kron(ed, np.ones((24, 1), np.int64)).ravel())
    assert np.array_equal(jK, np.kron(ed, np.ones((1, 24), np.int64)).ravel())
    f, fixed = s._get_bcs(2.0)
    fo, fixedo = O.get_bcs_cpu(s.bcs, 3, 2, 2, 2.0)
    assert np.array_equal(f, fo) and np.array_equal(np.sort(fixed), np.sort(fixedo))


def test_dense_solver_bcs():
    from genpipeline_amd.topology.simp_solver_dense_gpu import dense_bcs
    nx, ny, nz = 4, 3, 2
    f, fixed = dense_bcs(nx, ny, nz, 1000.0)
    i, j, k = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    sup = ((i == 0) | (j == 0) | (k == 0)).ravel()
    assert len(fixed) == 3 * sup.sum()
    load = (i == nx).ravel()
    share = -1000.0 / load.sum()
    fz = f[2::3]
    assert np.allclose(fz[load & ~sup], share) and np.all(fz[sup] == 0)
    assert np.all(f[0::3] == 0) and np.all(f[1::3] == 0)


def test_pipeline_utils():
    from genpipeline_amd.pipeline_utils import (FEM_SENTINEL, FEM_VALID_THRESHOLD, NumpyEncoder,
                                                is_valid_fem_result, smooth_voxels)
    from genpipeline_amd.schema import FEMResult
    assert FEM_SENTINEL == 1e6 and FEM_VALID_THRESHOLD == 1e5
    assert is_valid_fem_result({"stress": 12.0})
    assert is_valid_fem_result({"stress_max": 99999.0})
    assert not is_valid_fem_result({"stress": FEM_SENTINEL})
    assert not is_valid_fem_result({"stress": 0.0})
    assert not is_valid_fem_result({})
    assert not is_valid_fem_result(None)
    assert is_valid_fem_result(FEMResult(stress_max=5.0, compliance=1.0, mass=1.0))
    # constant field: blur keeps it, logistic maps 0.5 -> 0.5 and 0.6 -> 1/(1+e^-1.5)
    v = smooth_voxels(np.full((4, 4, 4), 0.5))
    assert np.allclose(v, 0.5)
    v = smooth_voxels(np.full((4, 4, 4), 0.6))
    assert np.allclose(v, 1.0 / (1.0 + np.exp(-1.5)))
    rnd = smooth_voxels(np.random.default_rng(0).uniform(size=(6, 5, 4)))
    assert rnd.shape == (6, 5, 4) and np.all((rnd > 0) & (rnd < 1))
    s = json.dumps({"a": np.arange(3), "b": np.float64(1.5), "c": np.int64(2),
                    "r": FEMResult(stress_max=1.0, compliance=2.0, mass=3.0)}, cls=NumpyEncoder)
    d = json.loads(s)
    assert d["a"] == [0, 1, 2] and d["r"]["compliance"] == 2.0


class _TinyDecoder:
    """Stands in for the VAE: decode() returns a fixed 16^3 field."""

    def __init__(self):
        import torch
        self._p = torch.nn.Parameter(torch.zeros(1))
        g = np.zeros((16, 16, 16), np.float32)
        g[:, 4:12, 4:12] = 1.0
        self.field = torch.from_numpy(g)

    def parameters(self):
        return iter([self._p])

    def decode(self, z):
        return self.field[None, None].expand(z.shape[0], 1, *self.field.shape)


def test_evaluator_penalty_path_without_gpu_solver(tmp_path):
    from genpipeline_amd.fem import VoxelFEMEvaluator
    from genpipeline_amd.pipeline_utils import FEM_SENTINEL
    from genpipeline_amd.schema import DesignParameters
    ev = VoxelFEMEvaluator(output_dir=str(tmp_path / "o"), vae_model=_TinyDecoder())
    r = ev.evaluate(np.zeros(8), ev.vae_model)
    assert r == {"stress": FEM_SENTINEL, "compliance": FEM_SENTINEL, "mass": 1.0,
                 "failure_reason": "no_ccx"}
    params = [DesignParameters(h_mm=1.0, r_mm=1.0, latent_z=[0.0] * 8),
              DesignParameters(h_mm=1.0, r_mm=1.0)]
    res = ev.evaluate_batch(params)
    assert [x.success for x in res] == [False, False]
    assert res[1].stress_max == FEM_SENTINEL
    assert len(ev.evaluation_history) == 1  # only entries with a latent vector
    ev.save_history(str(tmp_path / "h.json"))
    h = json.loads((tmp_path / "h.json").read_text())
    assert h[0]["result"]["stress_max"] == FEM_SENTINEL
    with pytest.raises(RuntimeError):
        VoxelFEMEvaluator().evaluate_batch(params)


def test_topo_data_generator_interface_and_draws(tmp_path):
    """Constructor defaults and the random draws of generate_single
    (topo_data_gen.py:129-264): geometry by weight (bridge, arch, 3dload
    twice as likely), grid by size class, volfrac and force in the organic
    ranges, all from NumPy's global generator in that order."""
    from genpipeline_amd.topology.topo_data_gen import TopoDataGenerator
    assert _defaults(TopoDataGenerator.__init__) == dict(
        output_dir="./fem_data", n_workers=2, use_gpu=True, n_iters=30, grid_size="medium",
        penal=3.5, organic_mode=True)
    assert _defaults(TopoDataGenerator.generate) == dict(n_samples=100)
    assert _defaults(TopoDataGenerator.generate_scaled) == dict(n_samples=1000, batch_size=100)
    g = TopoDataGenerator(output_dir=str(tmp_path / "d"))
    assert len(g.GEOM_BCS) == 17 and g.GEOM_BCS["spoke"]["fixed_face"] == "center"
    geoms = list(g.GEOM_BCS)
    w = np.array([g.GEOM_WEIGHTS.get(x, 1) for x in geoms], float)
    assert w.sum() == 20 and w[geoms.index("bridge")] == 2 and w[geoms.index("portal")] == 1
    np.random.seed(7)
    got = [g.draw_sample() for _ in range(40)]
    np.random.seed(7)
    for s in got:
        geom = np.random.choice(geoms, p=w / w.sum())
        assert s.geom == geom
        if geom in ("lbracket", "bridge", "arch", "center"):
            lo_hi = ((16, 24), (6, 10), (12, 18))
        elif geom in ("3dload", "offset"):
            lo_hi = ((22, 32), (8, 12), (8, 12))
        elif geom in ("branch", "y_tree", "network"):
            lo_hi = ((22, 32), (10, 16), (10, 16))
        else:
            lo_hi = ((20, 32), (6, 10), (6, 10))
        assert (s.nx, s.ny, s.nz) == tuple(np.random.randint(a, b) for a, b in lo_hi)
        assert s.volfrac == np.random.uniform(0.12, 0.35)
        assert s.force_mag == np.random.uniform(200, 5000)
    g.grid_size = "small"
    assert g._get_grid_dims("bridge") == (16, 8, 8)
