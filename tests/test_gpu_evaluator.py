"""Evaluator on the device: pre-processing parity with the reference chain and
the batched evaluate_batch against one-by-one evaluation.

The pre-processing oracle is the reference's own SciPy chain
(oracle/evaluator_oracle.py: zoom(0.5, order=1), gaussian_filter(0.5),
float32 logistic). The kernels follow SciPy's operation order and float32
rounding points, and are bit-equal to that chain when its logistic takes exp
in double rounded once (exp="double"), as the kernel does. Against NumPy's
own float32 exp (the literal reference; a CPU-dependent SIMD routine) voxels
differ by at most a few float32 ulps: tolerance 1e-6 absolute on values in
[0, 1].
"""
import numpy as np
import pytest
import torch

from oracle import evaluator_oracle as V

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,batch", [((64, 64, 64), 3), ((33, 20, 9), 2),
                                         ((64, 32, 16), 2), ((5, 4, 3), 1)])
def test_voxel_prep_matches_reference_chain(gpu, shape, batch):
    from genpipeline_amd.fem.voxel_prep import prep_voxels
    print(f"voxel_prep shape={shape} batch={batch}", flush=True)
    rng = np.random.default_rng(sum(shape))
    dec = rng.random((batch, 1, *shape), dtype=np.float32)
    out = prep_voxels(torch.from_numpy(dec).to(gpu)).cpu().numpy()
    for b in range(batch):
        ref = V.evaluator_voxels(dec[b, 0])
        assert out[b].shape == ref.shape
        d = np.abs(out[b].astype(np.float64) - ref.astype(np.float64))
        assert d.max() <= 1e-6, (b, d.max())
        assert np.array_equal(out[b], V.evaluator_voxels(dec[b, 0], exp="double"))


def test_voxel_prep_rejects_bad_input(gpu):
    from genpipeline_amd import _lib as L
    from genpipeline_amd.fem.voxel_prep import prep_voxels
    with pytest.raises(RuntimeError):
        prep_voxels(torch.zeros(1, 8, 8, 8))  # host tensor: no CPU fallback
    with pytest.raises(ValueError):
        prep_voxels(torch.zeros(1, 2, 8, 8, 8, device=gpu))
    with pytest.raises(L.GpError):
        prep_voxels(torch.zeros(1, 8, 8, 8, device=gpu), sigma=5.0)  # radius 20 > 16


def _evaluator(gpu, seed=0):
    from genpipeline_amd.fem import VoxelFEMEvaluator
    from genpipeline_amd.problems import SyntheticDecoder
    dec = SyntheticDecoder(seed=seed, device=gpu)
    return VoxelFEMEvaluator(vae_model=dec, use_gpu=True), dec


def test_evaluate_batch_matches_one_by_one(gpu):
    """Batched (FP64 refinement to 1e-8 around FP32 Jacobi-PCG inner solves,
    as scheduler jobs) against evaluate() (GPUVoxelFEM: FP64-refined to 1e-8)
    on the same latent vectors: compliance and stress within 1e-6 relative;
    and the reference chain on the host for the voxel fields."""
    from genpipeline_amd.problems import bo_latents
    from genpipeline_amd.schema import DesignParameters
    from genpipeline_amd.fem.voxel_fem import decode_batch
    ev, dec = _evaluator(gpu)
    zs = bo_latents(5, seed=11)
    params = [DesignParameters(h_mm=1, r_mm=1, latent_z=list(z)) for z in zs[:2]]
    params.append(DesignParameters(h_mm=1, r_mm=1))  # no latent vector
    params += [DesignParameters(h_mm=1, r_mm=1, latent_z=list(z)) for z in zs[2:]]
    print("evaluate_batch of 6 entries (5 latent vectors)", flush=True)
    res = ev.evaluate_batch(params)
    assert [r.success for r in res] == [True, True, False, True, True, True]
    assert len(ev.evaluation_history) == 10  # a result dict and a sample per latent vector
    raw = decode_batch(list(zs), dec).cpu().numpy()
    single, _ = _evaluator(gpu)
    fast, _ = _evaluator(gpu)
    fast.batch_refine = False  # one FP32 solve per field at 1e-6
    unrefined = fast.evaluate_latents(list(zs))
    for r, u, z, d in zip([res[i] for i in (0, 1, 3, 4, 5)], unrefined, zs, raw):
        one = single.evaluate(z, dec)
        vox = V.evaluator_voxels(d)
        mass = 32 ** 3 * float((vox > 0.5).mean()) * 7900 / 1e9
        assert r.mass == pytest.approx(mass, rel=1e-12)
        assert one["mass"] == pytest.approx(mass, rel=1e-12)
        rel = abs(r.compliance - one["compliance"]) / one["compliance"]
        print(f"compliance batch {r.compliance:.9e} single {one['compliance']:.9e} rel {rel:.2e}",
              flush=True)
        assert rel <= 1e-6
        assert r.stress_max == pytest.approx(one["stress"], rel=1e-6)
        # FP32 alone stops at its accuracy floor on solid/void fields
        assert u["compliance"] == pytest.approx(one["compliance"], rel=3e-5)
        assert u["mass"] == r.mass
    ev.close()
    single.close()
    fast.close()


def test_evaluate_batch_result_independent_of_batch(gpu):
    """A latent vector's result is bitwise the same alone and inside a batch."""
    from genpipeline_amd.problems import bo_latents
    ev, _ = _evaluator(gpu, seed=1)
    zs = bo_latents(6, seed=5)
    print("evaluate_latents alone vs batch of 6", flush=True)
    alone = ev.evaluate_latents([zs[3]])[0]
    batch = ev.evaluate_latents(list(zs))
    assert batch[3] == alone
    ev.close()


def test_sched_rhs_replaces_the_load(gpu):
    """A single solve admitted with rhs=g is bitwise the solve of a problem
    whose own load is g (same kernels, same order of operations)."""
    from genpipeline_amd.problems import point_load_vector
    from genpipeline_amd.sched import Scheduler
    from genpipeline_amd.serve import ServeConfig, make_problem
    shape = (32, 32, 32)
    cfg = ServeConfig()
    rng = np.random.default_rng(2)
    dens = rng.uniform(0.05, 1.0, 32 ** 3)
    P = make_problem(shape, "solve", cfg, gpu)
    g = point_load_vector(*shape, np.random.default_rng(3), 3)
    Q = make_problem(shape, "solve", cfg, gpu)
    from genpipeline_amd.problem import Problem
    Pg = Problem(Q.spec, Q.fixed, g, device=gpu)
    out = []
    for prob, rhs in ((P, torch.as_tensor(g, device=gpu)), (Pg, None)):
        S = Scheduler([(*shape, 8)], tile=(14, 18), device=gpu)
        S.admit(0, prob, n_iters=0, x0=dens, xphys0=dens, tol=1e-6, warm_start=False, rhs=rhs)
        _, st = S.run()
        out.append((S.jobs[0].u.clone(), st[0]))
        torch.cuda.synchronize()
        S.close()
    assert out[0][1].status == 0 and out[0][1].pcg_iters == out[1][1].pcg_iters
    assert out[0][1].compliance == out[1][1].compliance
    assert torch.equal(out[0][0], out[1][0])
    # a rejected admission leaves the slot free: the same slot then takes a
    # job and runs it to the same result
    S = Scheduler([(*shape, 8)], tile=(14, 18), device=gpu)
    try:
        with pytest.raises(ValueError):
            S.admit(0, P, n_iters=5, rhs=torch.as_tensor(g, device=gpu), warm_start=False)
        with pytest.raises(ValueError):
            S.admit(0, P, n_iters=0, rhs=torch.as_tensor(g[:-3], device=gpu), warm_start=False)
        assert S.jobs[0] is None
        S.admit(0, Pg, n_iters=0, x0=dens, xphys0=dens, tol=1e-6, warm_start=False)
        _, st = S.run()
        assert st[0].status == 0 and st[0].compliance == out[1][1].compliance
        assert torch.equal(S.jobs[0].u, out[1][0])
        torch.cuda.synchronize()
    finally:
        S.close()
    for p in (P, Q, Pg):
        p.close()
