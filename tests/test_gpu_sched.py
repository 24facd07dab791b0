"""Device scheduler: a job gives bitwise the results of Problem.run with the
same K.u tile, alone, in a mixed batch, admitted while other slots run, and
with ticks replayed from a HIP graph."""
import numpy as np
import pytest
import torch

from genpipeline_amd import Problem, ProblemSpec
from genpipeline_amd.bcs import clamp_x0
from genpipeline_amd.problems import point_load_vector
from oracle import simp_oracle as O

pytestmark = pytest.mark.gpu


def _problem(dev, nx, ny, nz, seed, volfrac=0.3, law="emin", oc="cpu", penal=3.0, pres=False):
    rng = np.random.default_rng(seed)
    f = point_load_vector(nx, ny, nz, rng, 2)
    spec = ProblemSpec(nx, ny, nz, volfrac=volfrac, law=law, oc=oc, penal=penal)
    mask = None
    if pres:
        mask = np.zeros((nx, ny, nz), bool)
        mask[nx - 2:, :, :2] = True
    return Problem(spec, clamp_x0(nx, ny, nz), f, preserved=mask, device=dev)


def _alone(P, tile, ni, n_iters, **kw):
    P.set_ku_tile(tile[0], tile[1], ni, precision="fp32")
    return P.run(n_iters, precision="fp32", **kw)


def _same(a, b):
    assert torch.equal(a.x, b.x) and torch.equal(a.xPhys, b.xPhys) and torch.equal(a.u, b.u)
    assert a.compliance == b.compliance
    assert a.pcg_iters == b.pcg_iters and a.bisect_steps == b.bisect_steps
    assert a.iters_done == b.iters_done and a.status == b.status


@pytest.mark.parametrize("tile", [(8, 32), (16, 16), (14, 18), (6, 34), (7, 34)])
def test_single_slot_equals_problem_run(gpu, tile):
    from genpipeline_amd.sched import Scheduler
    P = _problem(gpu, 12, 6, 5, 1, pres=True)
    ref = _alone(P, tile, 2, 4, tol=1e-6, max_iter=5000, warm_start=True)
    S = Scheduler([(12, 6, 5, 2)], tile=tile)
    S.admit(0, P, n_iters=4, tol=1e-6, max_iter=5000, warm_start=True)
    ticks, st = S.run()
    assert st[0].done and ticks > ref.pcg_iters
    _same(S.result(0, st[0]), ref)


def test_mixed_batch_each_equals_alone(gpu):
    from genpipeline_amd.sched import run_batch
    shapes = [(10, 5, 5), (16, 8, 8), (8, 8, 8), (20, 6, 6), (12, 10, 4)]
    Ps = [_problem(gpu, *s, seed=i, volfrac=0.25 + 0.03 * i,
                   oc="gpu" if i % 2 else "cpu", law="clamp" if i == 3 else "emin")
          for i, s in enumerate(shapes)]
    refs = [_alone(P, (8, 32), 2, 3, tol=1e-6, max_iter=5000, warm_start=True) for P in Ps]
    res, ticks = run_batch(Ps, 3, ni=2, graph_ticks=0, tol=1e-6, max_iter=5000)
    for a, b in zip(res, refs):
        _same(a, b)


def test_admission_while_running_and_graph_replay(gpu):
    """A job admitted into a free slot after other slots have run for a
    while, and ticks replayed from a captured graph (several replays with
    eager host work in between), still give the alone results."""
    from genpipeline_amd.sched import Scheduler
    A = _problem(gpu, 14, 6, 6, 3)
    B = _problem(gpu, 10, 8, 6, 4, volfrac=0.35)
    Cp = _problem(gpu, 12, 6, 6, 5, oc="gpu")
    ra = _alone(A, (8, 32), 2, 3, tol=1e-6, max_iter=5000)
    rb = _alone(B, (8, 32), 2, 2, tol=1e-6, max_iter=5000)
    rc = _alone(Cp, (8, 32), 2, 2, tol=1e-6, max_iter=5000)
    for graph in (0, 16):
        S = Scheduler([(16, 8, 8, 2)] * 3, graph_ticks=graph)
        S.admit(0, A, n_iters=3, tol=1e-6, max_iter=5000)
        S.tick(37)
        S.admit(1, B, n_iters=2, tol=1e-6, max_iter=5000)
        S.tick(64)
        st = S.status()
        assert not st[0].done and st[2].phase == 0
        S.admit(2, Cp, n_iters=2, tol=1e-6, max_iter=5000)
        for _ in range(3):  # replays interleaved with host-side checks
            S.tick(32)
            torch.cuda.synchronize()
            assert all(s.phase in range(0, 19) for s in S.status())
        _, st = S.run()
        assert all(s.done for s in st)
        _same(S.result(0, st[0]), ra)
        _same(S.result(1, st[1]), rb)
        _same(S.result(2, st[2]), rc)
        # a released slot can be refilled with a new job
        S.release(1)
        S.admit(1, B, n_iters=2, tol=1e-6, max_iter=5000)
        _, st = S.run()
        _same(S.result(1, st[1]), rb)
        S.close()


def test_gpu_solver_semantics_job(gpu):
    """Tolerance schedule, compliance solve after every update and early
    stop (SIMPSolverGPU.run) equal gp_simp_run with the same options."""
    from genpipeline_amd.sched import Scheduler
    P = _problem(gpu, 12, 6, 6, 7, law="clamp", oc="gpu", volfrac=0.35)
    kw = dict(tol=1e-6, max_iter=5000, tol_schedule=True, early_stop=True,
              compliance_each_iter=True, warm_start=True)
    ref = _alone(P, (8, 32), 2, 25, **kw)
    assert ref.iters_done < 25  # the early stop fired
    S = Scheduler([(12, 6, 6, 2)])
    S.admit(0, P, n_iters=25, **kw)
    _, st = S.run()
    _same(S.result(0, st[0]), ref)


def test_single_solve_job(gpu):
    """n_iters = 0: one solve at the given density; compliance f.u."""
    from genpipeline_amd.sched import Scheduler
    P = _problem(gpu, 10, 6, 6, 8, law="clamp", penal=1.0)
    dens = np.random.default_rng(2).uniform(0.01, 1.0, P.spec.n_elem)
    ref = _alone(P, (8, 32), 2, 0, x0=dens, xphys0=dens, tol=1e-6, max_iter=5000,
                 warm_start=False)
    S = Scheduler([(10, 6, 6, 2)])
    S.admit(0, P, n_iters=0, x0=dens, xphys0=dens, tol=1e-6, max_iter=5000, warm_start=False)
    _, st = S.run()
    _same(S.result(0, st[0]), ref)
    # and it is the linear-elastic compliance of that density field
    Po = O.OracleProblem(10, 6, 6, P.f_host, P.fixed, penal=1.0, law="clamp")
    u, _ = Po.solve_u(dens, "direct")
    assert abs(st[0].compliance - P.f_host @ u) <= 1e-5 * abs(P.f_host @ u)


def test_admission_errors(gpu):
    from genpipeline_amd import _lib
    from genpipeline_amd.sched import Scheduler
    P = _problem(gpu, 12, 6, 6, 9)
    S = Scheduler([(8, 6, 6, 2), (12, 6, 6, 2)])
    with pytest.raises(_lib.GpError, match="does not fit"):
        S.admit(0, P, n_iters=1)
    S.admit(1, P, n_iters=1)
    with pytest.raises(_lib.GpError, match="occupied"):
        S.admit(1, P, n_iters=1)
    with pytest.raises(_lib.GpError):
        Scheduler([(8, 8, 8, 2)], tile=(8, 32)).admit(5, P)
    S.run()


def test_compacted_launches_each_equals_alone(gpu):
    """Once the small jobs finish and a large one (>= 96 sweep blocks) runs
    on, ticks launch only the running jobs' blocks instead of replaying the
    graph of every block; a slot refilled meanwhile joins the list. Every
    job still equals Problem.run bit for bit."""
    from genpipeline_amd.sched import Scheduler
    big = _problem(gpu, 48, 24, 24, 7, volfrac=0.3)
    small = [_problem(gpu, *s, seed=10 + i) for i, s in enumerate([(8, 8, 8), (10, 6, 6), (12, 6, 5)])]
    late = _problem(gpu, 16, 8, 8, 21, oc="gpu")
    ref_big = _alone(big, (14, 18), 2, 4, tol=1e-6, max_iter=5000, warm_start=True)
    ref_small = [_alone(P, (14, 18), 2, 1, tol=1e-6, max_iter=5000, warm_start=True) for P in small]
    ref_late = _alone(late, (14, 18), 2, 2, tol=1e-6, max_iter=5000, warm_start=True)
    S = Scheduler([(48, 24, 24, 2), (8, 8, 8, 2), (10, 6, 6, 2), (16, 8, 8, 2)], tile=(14, 18),
                  graph_ticks=32)
    try:
        assert S.blocks(0) >= 96  # large enough for compacted launches
        S.admit(0, big, n_iters=4, tol=1e-6, max_iter=5000, warm_start=True)
        for k, P in enumerate(small[:2]):
            S.admit(1 + k, P, n_iters=1, tol=1e-6, max_iter=5000, warm_start=True)
        S.admit(3, small[2], n_iters=1, tol=1e-6, max_iter=5000, warm_start=True)
        print("compacted launches: running until the small jobs finish", flush=True)
        results, refilled, res_big, res_late = {}, False, None, None
        for _ in range(20000):
            S.tick(32)
            st = S.status()
            for slot in (1, 2):
                if slot not in results and st[slot].done:
                    results[slot] = S.result(slot, st[slot])
            if not refilled and st[3].done:
                _same(S.result(3, st[3]), ref_small[2])
                S.release(3)
                S.admit(3, late, n_iters=2, tol=1e-6, max_iter=5000, warm_start=True)
                refilled = True
                continue  # the next status read is the new job's
            if refilled and st[3].done and st[0].done:
                res_late, res_big = S.result(3, st[3]), S.result(0, st[0])
                break
        assert res_big is not None and res_late is not None
        _same(res_big, ref_big)
        _same(res_late, ref_late)
        _same(results[1], ref_small[0])
        _same(results[2], ref_small[1])
    finally:
        torch.cuda.synchronize()
        S.close()


def test_roller_supports_job_equals_alone(gpu):
    """A roller (one fixed component per node on face y = 0) beside the
    clamp: the scheduled job equals Problem.run bit for bit, and every fixed
    DOF of u stays exactly 0."""
    from genpipeline_amd.sched import Scheduler
    nx, ny, nz = 9, 7, 5
    rng = np.random.default_rng(41)
    f = point_load_vector(nx, ny, nz, rng, 2)
    roller = [3 * O.node_id(a, 0, k, ny, nz) + 2 for a in range(1, nx + 1) for k in range(nz + 1)]
    fixed = np.union1d(clamp_x0(nx, ny, nz), np.asarray(roller, np.int64))
    f[fixed] = 0.0
    P = Problem(ProblemSpec(nx, ny, nz, volfrac=0.3), fixed, f, device=gpu)
    for tile in ((14, 18), (8, 32)):
        ref = _alone(P, tile, 2, 4, tol=1e-6, max_iter=5000, warm_start=True)
        S = Scheduler([(nx, ny, nz, 2)], tile=tile)
        S.admit(0, P, n_iters=4, tol=1e-6, max_iter=5000, warm_start=True)
        _, st = S.run()
        res = S.result(0, st[0])
        _same(res, ref)
        assert torch.all(res.u[torch.from_numpy(fixed).to(gpu)] == 0)
        torch.cuda.synchronize()
        S.close()
    P.close()


@pytest.mark.parametrize("semantics", ["cpu_solver", "gpu_solver", "single_solve"])
def test_refined_job_equals_problem_run(gpu, semantics):
    """FP64 iterative refinement around FP32 inner solves as a scheduler job
    (residual, sums, inner solve, correction as phases) gives bitwise
    Problem.run(refine=True)'s results, for SIMPSolver semantics (warm
    starts), SIMPSolverGPU semantics (scheduled outer tolerance, compliance
    solve after every update, early stop) and a single solve; the refined
    answer is the FP64 oracle's to the outer tolerance."""
    from genpipeline_amd.sched import Scheduler
    if semantics == "cpu_solver":
        P = _problem(gpu, 12, 6, 5, 31, pres=True)
        kw = dict(tol=1e-10, inner_tol=1e-3, max_iter=5000, warm_start=True)
        n = 4
    elif semantics == "gpu_solver":
        P = _problem(gpu, 12, 6, 6, 32, law="clamp", oc="gpu", volfrac=0.35)
        kw = dict(tol=1e-6, inner_tol=1e-2, max_iter=5000, tol_schedule=True, early_stop=True,
                  compliance_each_iter=True, warm_start=True)
        n = 25
    else:
        P = _problem(gpu, 10, 6, 6, 33, law="clamp", penal=1.0)
        kw = dict(tol=1e-10, inner_tol=1e-3, max_iter=5000, warm_start=False)
        n = 0
    dens = np.random.default_rng(3).uniform(0.05, 1.0, P.spec.n_elem) if n == 0 else None
    x0 = dict(x0=dens, xphys0=dens) if dens is not None else {}
    P.set_ku_tile(14, 18, 2, precision="fp32")
    ref = P.run(n, precision="fp32", refine=True, **x0, **kw)
    if semantics == "gpu_solver":
        assert ref.iters_done < n  # the early stop fired
    assert ref.status == 0
    S = Scheduler([(P.spec.nx, P.spec.ny, P.spec.nz, 2)], tile=(14, 18))
    S.admit(0, P, n_iters=n, refine=True, **x0, **kw)
    _, st = S.run()
    _same(S.result(0, st[0]), ref)
    torch.cuda.synchronize()
    S.close()
    if n == 0:
        Po = O.OracleProblem(10, 6, 6, P.f_host, P.fixed, penal=1.0, law="clamp")
        u, _ = Po.solve_u(dens, "direct")
        assert np.abs(ref.u.cpu().numpy() - u).max() <= 1e-8 * np.abs(u).max()


def test_refined_and_fp32_jobs_share_ticks(gpu):
    """Refined and plain FP32 jobs of different sizes in one scheduler, one
    admitted while the others run: each equals its job alone."""
    from genpipeline_amd.sched import Scheduler
    A = _problem(gpu, 14, 6, 6, 41)
    B = _problem(gpu, 10, 8, 6, 42, volfrac=0.35, oc="gpu")
    Cp = _problem(gpu, 12, 6, 6, 43)
    tile = (14, 18)
    for P in (A, B, Cp):
        P.set_ku_tile(*tile, 2, precision="fp32")
    ka = dict(tol=1e-9, inner_tol=1e-3, max_iter=5000, refine=True)
    kb = dict(tol=1e-6, max_iter=5000)
    ra = A.run(3, precision="fp32", **ka)
    rb = B.run(3, precision="fp32", **kb)
    rc = Cp.run(2, precision="fp32", **ka)
    S = Scheduler([(16, 8, 8, 2)] * 3, tile=tile, graph_ticks=16)
    S.admit(0, A, n_iters=3, **ka)
    S.admit(1, B, n_iters=3, **kb)
    S.tick(50)
    S.admit(2, Cp, n_iters=2, **ka)
    _, st = S.run()
    _same(S.result(0, st[0]), ra)
    _same(S.result(1, st[1]), rb)
    _same(S.result(2, st[2]), rc)
    torch.cuda.synchronize()
    S.close()


def test_warm_compliance_job_equals_alone(gpu):
    """Compliance solves warm-started from the previous displacement (an
    option beside SIMPSolverGPU's zero start): the scheduled job equals
    Problem.run with the option bit for bit, runs fewer PCG iterations than
    the zero-start run, and ends within the tolerance of it."""
    from genpipeline_amd.sched import Scheduler
    P = _problem(gpu, 12, 6, 6, 7, law="clamp", oc="gpu", volfrac=0.35)
    kw = dict(tol=1e-6, max_iter=5000, tol_schedule=True, early_stop=True,
              compliance_each_iter=True, warm_start=True)
    cold = _alone(P, (8, 32), 2, 25, **kw)
    warm = _alone(P, (8, 32), 2, 25, warm_compliance=True, **kw)
    assert warm.pcg_iters < cold.pcg_iters
    assert abs(warm.compliance - cold.compliance) <= 1e-4 * abs(cold.compliance)
    S = Scheduler([(12, 6, 6, 2)])
    S.admit(0, P, n_iters=25, warm_compliance=True, **kw)
    _, st = S.run()
    _same(S.result(0, st[0]), warm)
    torch.cuda.synchronize()
    S.close()


def test_autotuned_shape_batch_equals_alone(gpu):
    """The autotuner picks the scheduler's tile and a sweep depth per grid size
    from timed trial runs of the batch; a batch run with tile="auto" then uses
    that shape, and every job is bitwise Problem.run with the tile and depth
    the autotuner gave its grid."""
    from genpipeline_amd import autotune
    from genpipeline_amd.sched import run_batch
    shapes = [(16, 8, 8), (10, 6, 6), (16, 8, 8), (12, 6, 5)]
    Ps = [_problem(gpu, *s, seed=20 + i, volfrac=0.3) for i, s in enumerate(shapes)]
    tile, depth, timings = autotune.tune_sched(Ps, n_iters=1, tiles=[(8, 32), (14, 18)],
                                               nis=(2, 4), persist=False, max_iter=5000)
    grids = {tuple(s) for s in shapes}
    assert tile in ((8, 32), (14, 18)) and set(depth) == grids
    assert all(depth[g] in (2, 4, autotune.SCHED_START_NI) for g in grids)
    # two tiles at the start depth, then up to two depths per grid size
    assert 2 <= len(timings) <= 2 + 2 * len(grids)
    assert all(t > 0 for t in timings.values())
    assert autotune.lookup_sched(shapes, gpu) == (tile, depth)
    res, _ = run_batch(Ps, 2, tile="auto", graph_ticks=0, tol=1e-6, max_iter=5000)
    for P, s, r in zip(Ps, shapes, res):
        _same(r, _alone(P, tile, depth[tuple(s)], 2, tol=1e-6, max_iter=5000))
