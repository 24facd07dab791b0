"""Batched execution: every problem's result in a mixed batch is bitwise the
same as the problem solved alone with the same K.u tile, and matches the FP64
oracle to the FP32 solve tolerance."""
import numpy as np
import pytest
import torch

from oracle import simp_oracle as O
from genpipeline_amd import Problem, ProblemSpec
from genpipeline_amd.batch import Batch
from genpipeline_amd.bcs import clamp_x0
from genpipeline_amd.problems import point_load_vector

pytestmark = pytest.mark.gpu

SIZES = [(12, 6, 6), (9, 5, 7), (16, 8, 8), (6, 6, 6), (11, 4, 9)]


def _make(gpu, seed, grid, oc="cpu", law="emin"):
    rng = np.random.default_rng(seed)
    vf = float(rng.uniform(0.25, 0.4))
    f = point_load_vector(*grid, rng, int(rng.integers(1, 4)))
    spec = ProblemSpec(*grid, volfrac=vf, oc=oc, law=law)
    return Problem(spec, clamp_x0(*grid), f, device=gpu), vf, f


@pytest.mark.parametrize("tile", [(8, 32), (6, 34)])
def test_batch_equals_alone_bitwise(gpu, tile):
    made = [_make(gpu, 100 + i, g, oc="cpu" if i % 2 else "gpu") for i, g in enumerate(SIZES)]
    probs = [m[0] for m in made]
    with Batch(probs, tile=tile) as B:
        res_b = B.run(6, volfracs=[m[1] for m in made], tol=1e-6)
    for i, (P, vf, f) in enumerate(made):
        r = P.run(6, volfrac=vf, tol=1e-6, precision="fp32", warm_start=True)
        assert r.compliance == res_b[i].compliance, i
        assert torch.equal(r.xPhys, res_b[i].xPhys), i
        assert r.pcg_iters == res_b[i].pcg_iters, i
        P.close()


def test_batch_order_invariance(gpu):
    made = [_make(gpu, 200 + i, g) for i, g in enumerate(SIZES[:4])]
    probs = [m[0] for m in made]
    vfs = [m[1] for m in made]
    with Batch(probs) as B:
        r1 = B.run(4, volfracs=vfs)
    with Batch(probs[::-1]) as B:
        r2 = B.run(4, volfracs=vfs[::-1])[::-1]
    for a, b in zip(r1, r2):
        assert a.compliance == b.compliance and torch.equal(a.xPhys, b.xPhys)
    for p in probs:
        p.close()


def test_batch_matches_oracle(gpu):
    grid = (12, 6, 6)
    P, vf, f = _make(gpu, 7, grid)
    P2, vf2, f2 = _make(gpu, 8, (8, 4, 4))
    with Batch([P, P2]) as B:
        res = B.run(5, volfracs=[vf, vf2], tol=1e-8)
    ref = O.OracleProblem(*grid, f=f, fixed=clamp_x0(*grid), volfrac=vf)
    xp, c, _ = ref.run(5, "pcg", 1e-12)
    assert abs(res[0].compliance - c) <= 1e-5 * abs(c)
    assert np.abs(res[0].xPhys.cpu().numpy() - xp).max() <= 1e-4
    P.close()
    P2.close()


def test_batch_with_roller_supports_equals_alone(gpu):
    """Supports that fix one component of a node (a roller on face y = 0
    beside the clamp): the lockstep batch reads one Jacobi value per node as
    the single-problem sweep does, gives bitwise Problem.run's results, and
    every fixed DOF of u stays exactly 0."""
    probs, vfs, fixeds = [], [], []
    for i, (nx, ny, nz) in enumerate([(9, 7, 5), (12, 6, 6)]):
        rng = np.random.default_rng(300 + i)
        vf = float(rng.uniform(0.25, 0.4))
        f = point_load_vector(nx, ny, nz, rng, 2)
        roller = [3 * O.node_id(a, 0, k, ny, nz) + 2 for a in range(1, nx + 1) for k in range(nz + 1)]
        fixed = np.union1d(clamp_x0(nx, ny, nz), np.asarray(roller, np.int64))
        f[fixed] = 0.0
        probs.append(Problem(ProblemSpec(nx, ny, nz, volfrac=vf), fixed, f, device=gpu))
        vfs.append(vf)
        fixeds.append(fixed)
    with Batch(probs) as B:
        res_b = B.run(4, volfracs=vfs, tol=1e-6)
    for P, vf, fixed, rb in zip(probs, vfs, fixeds, res_b):
        r = P.run(4, volfrac=vf, tol=1e-6, precision="fp32", warm_start=True)
        assert r.compliance == rb.compliance and r.pcg_iters == rb.pcg_iters
        assert torch.equal(r.xPhys, rb.xPhys) and torch.equal(r.u, rb.u)
        assert torch.all(rb.u[torch.from_numpy(fixed).to(gpu)] == 0)
        P.close()
