"""The single-problem PCG updates x and p two iterations at a time (odd
launches leave them, even launches apply both updates, an odd last iteration
is completed after the solve); the scheduler's sweep updates them every
iteration. Stopped by the iteration cap at odd and even counts, or by
convergence, both give the same bits."""
import numpy as np
import pytest
import torch

from genpipeline_amd import Problem, ProblemSpec
from genpipeline_amd.bcs import clamp_x0
from genpipeline_amd.problems import point_load_vector

pytestmark = pytest.mark.gpu


def _setup(gpu, seed=3):
    nx, ny, nz = 12, 6, 5
    rng = np.random.default_rng(seed)
    f = point_load_vector(nx, ny, nz, rng, 2)
    P = Problem(ProblemSpec(nx, ny, nz, volfrac=0.3), clamp_x0(nx, ny, nz), f, device=gpu)
    xphys = torch.from_numpy(rng.uniform(0.05, 1.0, P.spec.n_elem)).to(gpu)
    P.set_ku_tile(8, 32, 2, precision="fp32")
    return P, xphys


def _sched_solve(gpu, P, xphys, tol, max_iter):
    from genpipeline_amd.sched import Scheduler
    S = Scheduler([(P.spec.nx, P.spec.ny, P.spec.nz, 2)], tile=(8, 32))
    try:
        S.admit(0, P, n_iters=0, xphys0=xphys, tol=tol, max_iter=max_iter, warm_start=False)
        ticks, st = S.run()
        return S.result(0, st[0])
    finally:
        S.close()


@pytest.mark.parametrize("iters", [1, 2, 3, 4, 7, 8])
def test_capped_solve_bitwise_every_iteration_form(gpu, iters):
    P, xphys = _setup(gpu)
    E = P.stiffness(xphys)
    u, r = P.pcg(E, tol=0.0, max_iter=iters, precision="fp32")
    assert r.iters == iters
    ref = _sched_solve(gpu, P, xphys, 0.0, iters)
    assert ref.pcg_iters == iters
    assert torch.equal(ref.u, u.double())
    P.close()


@pytest.mark.parametrize("seed", [3, 4, 5, 6])
def test_converged_solve_bitwise_every_iteration_form(gpu, seed):
    P, xphys = _setup(gpu, seed)
    E = P.stiffness(xphys)
    u, r = P.pcg(E, tol=1e-6, max_iter=5000, precision="fp32")
    assert r.status == 0
    ref = _sched_solve(gpu, P, xphys, 1e-6, 5000)
    assert ref.pcg_iters == r.iters
    assert torch.equal(ref.u, u.double())
    P.close()
