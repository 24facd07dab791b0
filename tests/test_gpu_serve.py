"""Continuous-batching server: every request's result is the same alone and
in the mix (bitwise), admission respects the HBM budget, and the served
outputs follow the reference formulas."""
import numpy as np
import pytest
import torch

from oracle import simp_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = ((12, 6, 6), (16, 8, 8))


def _requests(n, lam=400.0, seed=3):
    from genpipeline_amd.problems import serving_requests
    return serving_requests(lam, n, seed=seed, p_small=0.6, p_solve=0.5, shapes=SHAPES)


def _same(a, b):
    assert a.compliance == b.compliance, (a.rid, a.compliance, b.compliance)
    assert a.stress_max == b.stress_max and a.displacement_max == b.displacement_max
    assert a.iters_done == b.iters_done and a.pcg_iters == b.pcg_iters and a.status == b.status
    if a.kind == "simp":
        assert torch.equal(a.xPhys, b.xPhys)


@pytest.mark.parametrize("policy,recycle", [("fifo", 0), ("sjf", 0), ("sjf", 4)])
def test_each_result_same_alone_and_in_mix(gpu, policy, recycle):
    """recycle: SIMP requests start each warm OC solve from the Galerkin
    combination of their last solutions (pooled problems carry the ring
    from request to request; each request's history starts empty)."""
    from genpipeline_amd.serve import Server, ServeConfig, run_alone
    cfg = ServeConfig(simp_iters=6, policy=policy, simp_recycle=recycle)
    reqs = _requests(14)
    assert {r.kind for r in reqs} == {"solve", "simp"} and {r.shape for r in reqs} == set(SHAPES)
    srv = Server({SHAPES[0]: 3, SHAPES[1]: 2}, cfg=cfg, tile=(16, 16), ni=4, chunk_ticks=32)
    try:
        res = srv.run(reqs, time_scale=1.0, keep_xphys=True)
    finally:
        torch.cuda.synchronize()
    # the requests overlapped on the device
    overlapping = sum(1 for a in res for b in res
                      if a.rid < b.rid and a.t_admit < b.t_done and b.t_admit < a.t_done)
    assert overlapping > 0
    for r, got in zip(reqs, res):
        assert got.rid == r.rid and got.t_done >= got.t_admit >= got.t_arrive - 1e-9
        _same(got, run_alone(r, cfg, tile=(16, 16), ni=4))
    srv.close()


def test_admission_respects_hbm_budget(gpu):
    from genpipeline_amd.problem import footprint_bytes
    from genpipeline_amd.serve import Server, ServeConfig, _spec, run_alone
    cfg = ServeConfig(simp_iters=3)
    reqs = _requests(8, lam=2000.0, seed=5)
    big = max(footprint_bytes(_spec(s, k)) for s in SHAPES for k in ("solve", "simp"))
    budget = 2 * big  # at most two problems resident at once
    srv = Server({SHAPES[0]: 4, SHAPES[1]: 4}, cfg=cfg, hbm_budget=budget, chunk_ticks=16)
    res = srv.run(reqs, keep_xphys=True)
    assert srv.peak_used <= budget
    for r, got in zip(reqs, res):
        _same(got, run_alone(r, cfg))
    srv.close()
    with pytest.raises(RuntimeError, match="HBM budget"):
        Server({reqs[0].shape: 1}, cfg=cfg, hbm_budget=1000).run(reqs[:1])


@pytest.mark.parametrize("refine", [True, False])
def test_served_solve_outputs_follow_reference_formulas(gpu, refine):
    """A "solve" request reports GPUVoxelFEM's outputs for its density
    (gpu_fem_solver.py:96-133). Served solves are FP64-refined to 1e-8
    around FP32 inner solves: compliance, stress and displacement within
    1e-6 of the FP64 direct solve. A plain FP32 solve at 1e-6 on this
    1e-3-contrast field stops at its accuracy floor: 1e-3."""
    from genpipeline_amd.bcs import simp_gpu_bcs
    from genpipeline_amd.serve import ServeConfig, run_alone
    # compliance ~2e2: inside every clip of the formulas
    cfg = ServeConfig(force_n=1.0, solve_refine=refine, solve_tol=1e-8 if refine else 1e-6)
    req = next(r for r in _requests(20, seed=11) if r.kind == "solve" and r.shape == SHAPES[1])
    got = run_alone(req, cfg)
    nx, ny, nz = req.shape
    f, fixed = simp_gpu_bcs(cfg.bcs, nx, ny, nz, cfg.force_n)
    f = f.astype(np.float32).astype(np.float64)
    P = O.OracleProblem(nx, ny, nz, f, fixed, penal=1.0, law="clamp")
    u, _ = P.solve_u(req.density, "direct")
    c = float(f @ u)
    n_solid = int((req.density > 0.5).sum())
    assert 1.0 < c < 1e8 and n_solid >= 50
    tol = 1e-6 if refine else 1e-3
    assert abs(got.compliance - c) <= tol * c
    assert abs(got.stress_max - min(10.0 * c / n_solid, 1e4)) <= tol * got.stress_max
    dmax = float(np.linalg.norm(u.reshape(-1, 3), axis=1).max())
    assert abs(got.displacement_max - min(dmax, 1e4)) <= tol * min(dmax, 1e4)
