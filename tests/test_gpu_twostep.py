"""The single-problem PCG and the device scheduler update x and p two
iterations at a time (odd launches leave them, even launches apply both
updates, an odd last iteration is completed after the solve); the scheduler
starts its launch 0 in the even form with zero history where the single
problem uses the every-iteration form. The lockstep batch updates them every
iteration. Stopped by the iteration cap at odd and even counts, or by
convergence, all give the same bits."""
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
def test_capped_solve_bitwise_scheduler(gpu, iters):
    P, xphys = _setup(gpu)
    E = P.stiffness(xphys)
    u, r = P.pcg(E, tol=0.0, max_iter=iters, precision="fp32")
    assert r.iters == iters
    ref = _sched_solve(gpu, P, xphys, 0.0, iters)
    assert ref.pcg_iters == iters
    assert torch.equal(ref.u, u.double())
    P.close()


@pytest.mark.parametrize("seed", [3, 4, 5, 6])
def test_converged_solve_bitwise_scheduler(gpu, seed):
    P, xphys = _setup(gpu, seed)
    E = P.stiffness(xphys)
    u, r = P.pcg(E, tol=1e-6, max_iter=5000, precision="fp32")
    assert r.status == 0
    ref = _sched_solve(gpu, P, xphys, 1e-6, 5000)
    assert ref.pcg_iters == r.iters
    assert torch.equal(ref.u, u.double())
    P.close()


@pytest.mark.parametrize("iters", [1, 2, 3, 7, 8])
def test_capped_run_bitwise_lockstep_every_iteration_form(gpu, iters):
    """Two OC iterations with every solve capped at `iters` PCG iterations:
    the lockstep batch's every-iteration x and p update and Problem.run's
    two-step update give the same u, density and compliance."""
    from genpipeline_amd.batch import Batch
    P, _ = _setup(gpu)
    ref = P.run(2, tol=0.0, max_iter=iters, precision="fp32", warm_start=True)
    assert ref.pcg_iters == 3 * iters  # two OC solves and the final one
    with Batch([P], tile=(8, 32)) as B:
        (got,) = B.run(2, tol=0.0, max_iter=iters)
    assert torch.equal(got.u, ref.u) and torch.equal(got.xPhys, ref.xPhys)
    assert got.compliance == ref.compliance and got.pcg_iters == ref.pcg_iters
    P.close()
