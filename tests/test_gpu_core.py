"""GPU parity of each hot-path kernel against the FP64 oracle on seeded inputs.

Tolerances: integer work bit-exact; FP64 kernels 1e-12 relative; FP32 K.u
2e-6 relative to max|y|; filter and OC bit-exact against the oracle's CSR
arithmetic where the operation order is the same.
"""
import numpy as np
import pytest
import torch

from oracle import simp_oracle as O
from genpipeline_amd import Problem, ProblemSpec
from genpipeline_amd.bcs import clamp_x0

pytestmark = pytest.mark.gpu

GRIDS = [(1, 1, 1), (3, 2, 4), (8, 5, 6), (16, 7, 9)]


def _problem(nx, ny, nz, law="emin", seed=0, **kw):
    rng = np.random.default_rng(seed)
    spec = ProblemSpec(nx, ny, nz, law=law, **kw)
    fixed = clamp_x0(nx, ny, nz)
    f = np.zeros(spec.n_dof)
    f[3 * (spec.n_nodes - 1) + 2] = -1.0
    f[3 * (spec.n_nodes - 1 - (nz + 1)) + 1] = 0.5
    x = rng.uniform(0.05, 1.0, spec.n_elem)
    return spec, fixed, f, x, rng


@pytest.mark.parametrize("grid", GRIDS)
def test_ku_fp64_matches_oracle(gpu, grid):
    spec, fixed, f, x, rng = _problem(*grid)
    KE = O.ke_hex8(0.3)
    edof = O.edof_matrix(*grid)
    E = O.stiffness(x, "emin", 3.0).astype(np.float32)
    u = rng.standard_normal(spec.n_dof)
    ref = O.ku_matfree(KE, E.astype(np.float64), u, edof, fixed)
    with Problem(spec, fixed, f, device=gpu) as P:
        y = P.ku(torch.from_numpy(E), torch.from_numpy(u)).cpu().numpy()
        y_nomask = P.ku(torch.from_numpy(E), torch.from_numpy(u), mask=False).cpu().numpy()
    scale = np.abs(ref).max()
    assert np.abs(y - ref).max() <= 1e-12 * scale
    ref_nm = O.ku_matfree(KE, E.astype(np.float64), u, edof)
    assert np.abs(y_nomask - ref_nm).max() <= 1e-12 * np.abs(ref_nm).max()


@pytest.mark.parametrize("grid", GRIDS)
def test_ku_fp32_matches_oracle(gpu, grid):
    spec, fixed, f, x, rng = _problem(*grid, seed=1)
    KE = O.ke_hex8(0.3)
    edof = O.edof_matrix(*grid)
    E = O.stiffness(x, "emin", 3.0).astype(np.float32)
    u = rng.standard_normal(spec.n_dof).astype(np.float32)
    ref = O.ku_matfree(KE, E.astype(np.float64), u.astype(np.float64), edof, fixed)
    with Problem(spec, fixed, f, device=gpu) as P:
        y = P.ku(torch.from_numpy(E), torch.from_numpy(u)).cpu().numpy().astype(np.float64)
    assert np.abs(y - ref).max() <= 2e-6 * np.abs(ref).max()


@pytest.mark.parametrize("tile", [(8, 32, 1), (16, 16, 3), (14, 18, 2), (6, 34, 5), (14, 34, 4),
                                  (4, 64, 7), (10, 26, 16), (0, 0, 0)])
@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_ku_every_tile_matches_oracle(gpu, tile, prec):
    """Modal sweep for each workgroup shape (tj=0: dense gather) on ragged grids."""
    KE = O.ke_hex8(0.3)
    for grid in [(5, 7, 6), (13, 4, 17), (2, 33, 9)]:
        spec, fixed, f, x, rng = _problem(*grid, seed=21)
        edof = O.edof_matrix(*grid)
        E = O.stiffness(x, "emin", 3.0).astype(np.float32)
        dt = np.float32 if prec == "fp32" else np.float64
        u = rng.standard_normal(spec.n_dof).astype(dt)
        ref = O.ku_matfree(KE, E.astype(np.float64), u.astype(np.float64), edof, fixed)
        with Problem(spec, fixed, f, device=gpu) as P:
            P.set_ku_tile(*tile, precision=prec)
            tj, tk, ni, modal = P.ku_tile(prec)
            assert modal == 1 and (tj, tk, ni) == tile
            y = P.ku(torch.from_numpy(E), torch.from_numpy(u)).cpu().numpy().astype(np.float64)
        tol = 2e-6 if prec == "fp32" else 1e-12
        assert np.abs(y - ref).max() <= tol * np.abs(ref).max(), (grid, tile)


def test_custom_table_uses_gather(gpu):
    spec, fixed, f, x, rng = _problem(4, 3, 2, seed=2)
    with Problem(spec, fixed, f, ke=rng.standard_normal((24, 24)), device=gpu) as P:
        assert P.ku_tile()[3] == 0 and P.ku_tile()[0] == 0


def test_custom_table_with_unequal_float_diagonals_uses_gather(gpu):
    """The float sweeps read one Jacobi value per node (the x diagonal). A
    custom table that matches the modal form to 1e-10 but whose float32
    diagonal entries of one corner differ keeps the per-DOF gather path, and
    still solves to the direct answer."""
    spec, fixed, f, x, rng = _problem(6, 3, 4, seed=4)
    KE = O.ke_hex8(0.3)
    # scale so the diagonal sits just below a float32 rounding midpoint, then
    # push one entry across it by 3e-11 relative
    s = (1.0 + 2.0 ** -24 - 1e-11) / KE[0, 0]
    base = KE * s
    bent = base.copy()
    bent[4, 4] *= 1.0 + 3e-11
    assert np.float32(base[4, 4]) != np.float32(bent[4, 4])
    with Problem(spec, fixed, f, ke=base, device=gpu) as P:
        assert P.ku_tile()[3] == 1
    E = O.stiffness(x, "emin", 3.0).astype(np.float32)
    K = O.assemble_K(bent.T.copy(), E.astype(np.float64), O.edof_matrix(6, 3, 4), spec.n_dof)
    ud, _ = O.solve(K, f, fixed, "direct")
    with Problem(spec, fixed, f, ke=bent, device=gpu) as P:
        assert P.ku_tile()[3] == 0 and P.ku_tile()[0] == 0
        u, st = P.pcg(torch.from_numpy(E), tol=1e-10, refine=True)
    assert st.status == 0
    assert np.abs(u.cpu().numpy() - ud).max() <= 1e-8 * np.abs(ud).max()


def test_ku_custom_table(gpu):
    """A caller-supplied (asymmetric) table is applied as y_e = KE u_e."""
    grid = (4, 3, 2)
    spec, fixed, f, x, rng = _problem(*grid, seed=2)
    KE = rng.standard_normal((24, 24))
    edof = O.edof_matrix(*grid)
    E = O.stiffness(x, "emin", 3.0).astype(np.float32)
    u = rng.standard_normal(spec.n_dof)
    # oracle ku_matfree applies KE^T per element (CPU assembly convention)
    ref = O.ku_matfree(KE.T.copy(), E.astype(np.float64), u, edof, fixed)
    with Problem(spec, fixed, f, ke=KE, device=gpu) as P:
        y = P.ku(torch.from_numpy(E), torch.from_numpy(u)).cpu().numpy()
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("grid", GRIDS[1:])
def test_diag_matches_assembled(gpu, grid):
    spec, fixed, f, x, rng = _problem(*grid, seed=3)
    KE = O.ke_hex8(0.3)
    edof = O.edof_matrix(*grid)
    E = O.stiffness(x, "emin", 3.0).astype(np.float32)
    K = O.assemble_K(KE, E.astype(np.float64), edof, spec.n_dof)
    with Problem(spec, fixed, f, device=gpu) as P:
        d = P.diag(torch.from_numpy(E)).cpu().numpy()
    ref = K.diagonal()
    assert np.abs(d - ref).max() <= 1e-13 * ref.max()


@pytest.mark.parametrize("kernel", ["modal", "gather"])
@pytest.mark.parametrize("precision,refine,tol,bound", [
    ("fp64", False, 1e-11, 1e-9),
    ("fp32", True, 1e-11, 1e-9),
    ("fp32", False, 1e-6, 2e-4),
])
def test_pcg_matches_direct(gpu, precision, refine, tol, bound, kernel):
    grid = (12, 6, 6)
    spec, fixed, f, x, rng = _problem(*grid, seed=4)
    KE = O.ke_hex8(0.3)
    edof = O.edof_matrix(*grid)
    E = O.stiffness(x, "emin", 3.0).astype(np.float32)
    K = O.assemble_K(KE, E.astype(np.float64), edof, spec.n_dof)
    u_ref, _ = O.solve(K, f, fixed, "direct")
    with Problem(spec, fixed, f, device=gpu) as P:
        if kernel == "gather":
            P.set_ku_tile(0, precision="fp32")
            P.set_ku_tile(0, precision="fp64")
        u, res = P.pcg(torch.from_numpy(E), tol=tol, max_iter=20000, precision=precision,
                       refine=refine, inner_tol=1e-4)
        u = u.cpu().numpy().astype(np.float64)
    assert res.status == 0, res
    assert np.abs(u - u_ref).max() <= bound * np.abs(u_ref).max()
    c, c_ref = f @ u, f @ u_ref
    assert abs(c - c_ref) <= max(bound, 1e-12) * abs(c_ref)
    assert np.all(u[fixed] == 0.0)


def test_pcg_warm_start_and_zero_load(gpu):
    grid = (6, 4, 4)
    spec, fixed, f, x, rng = _problem(*grid, seed=5)
    E = O.stiffness(x, "emin", 3.0).astype(np.float32)
    with Problem(spec, fixed, f, device=gpu) as P:
        u1, r1 = P.pcg(torch.from_numpy(E), tol=1e-10, precision="fp64")
        u2, r2 = P.pcg(torch.from_numpy(E), u0=u1, tol=1e-10, precision="fp64")
        assert r2.iters <= 1
        assert torch.allclose(u1, u2, rtol=0, atol=1e-12 * float(u1.abs().max()))
        uz, rz = P.pcg(torch.from_numpy(E), f=np.zeros(spec.n_dof), tol=1e-8, precision="fp32")
        assert rz.iters == 0 and float(uz.abs().max()) == 0.0


@pytest.mark.parametrize("kernel", ["modal", "gather"])
def test_pcg_max_iter_reports_not_converged(gpu, kernel):
    grid = (10, 5, 5)
    spec, fixed, f, x, rng = _problem(*grid, seed=6)
    E = O.stiffness(x, "emin", 3.0).astype(np.float32)
    with Problem(spec, fixed, f, device=gpu) as P:
        if kernel == "gather":
            P.set_ku_tile(0, precision="fp64")
        _, res = P.pcg(torch.from_numpy(E), tol=1e-14, max_iter=5, precision="fp64")
    assert res.status == 5 and res.iters == 5


@pytest.mark.parametrize("kernel", ["modal", "gather"])
@pytest.mark.parametrize("refine,tol,bound", [(True, 1e-11, 1e-9), (False, 1e-6, 2e-4)])
def test_fp32_pcg_with_partial_supports(gpu, kernel, refine, tol, bound):
    """Supports that fix only one component of a node (a roller on face
    y = 0 beside the clamp at x = 0). The float single-kernel sweep reads one
    Jacobi value per node and keeps the support bits out of the sweep; its
    solution matches the direct solve as the gather path's does, and every
    fixed DOF stays exactly 0."""
    grid = (9, 7, 5)
    nx, ny, nz = grid
    spec, fixed, f, x, rng = _problem(*grid, seed=21)
    roller = [3 * O.node_id(i, 0, k, ny, nz) + 2 for i in range(1, nx + 1) for k in range(nz + 1)]
    fixed = np.union1d(fixed, np.asarray(roller, dtype=np.int64))
    KE = O.ke_hex8(0.3)
    edof = O.edof_matrix(*grid)
    E = O.stiffness(x, "emin", 3.0).astype(np.float32)
    K = O.assemble_K(KE, E.astype(np.float64), edof, spec.n_dof)
    u_ref, _ = O.solve(K, f, fixed, "direct")
    with Problem(spec, fixed, f, device=gpu) as P:
        if kernel == "gather":
            P.set_ku_tile(0, precision="fp32")
        u, res = P.pcg(torch.from_numpy(E), tol=tol, max_iter=20000, precision="fp32",
                       refine=refine, inner_tol=1e-4)
        u = u.cpu().numpy().astype(np.float64)
    assert res.status == 0, res
    assert np.all(u[fixed] == 0.0)
    assert np.abs(u - u_ref).max() <= bound * np.abs(u_ref).max()


def test_pcg_modal_and_gather_agree(gpu):
    """Same iteration count and solution (FP64) with either K.u kernel."""
    grid = (9, 7, 5)
    spec, fixed, f, x, rng = _problem(*grid, seed=13)
    E = torch.from_numpy(O.stiffness(x, "emin", 3.0).astype(np.float32))
    with Problem(spec, fixed, f, device=gpu) as P:
        u1, r1 = P.pcg(E, tol=1e-10, precision="fp64")
        P.set_ku_tile(0, precision="fp64")
        u2, r2 = P.pcg(E, tol=1e-10, precision="fp64")
    assert abs(r1.iters - r2.iters) <= 1
    assert float((u1 - u2).abs().max()) <= 1e-9 * float(u1.abs().max())


def test_pcg_breakdown_on_indefinite_table(gpu):
    """The reference's own element table (simp_solver.py:284-1543) is
    asymmetric and its symmetric part has two negative eigenvalues: PCG must
    report breakdown (p^T A p <= 0) instead of returning garbage."""
    grid = (8, 4, 4)
    spec, fixed, f, x, rng = _problem(*grid, seed=7, law="clamp")
    KEr = O.ke_ref()
    E = O.stiffness(np.full(spec.n_elem, 0.4), "clamp", 3.0).astype(np.float32)
    with Problem(spec, fixed, f, ke=KEr.T, device=gpu) as P:
        _, res = P.pcg(torch.from_numpy(E), tol=1e-10, max_iter=5000, precision="fp64")
    assert res.status == 4


@pytest.mark.parametrize("law", ["emin", "clamp"])
@pytest.mark.parametrize("udtype", [np.float64, np.float32])
def test_sens_matches_oracle(gpu, law, udtype):
    grid = (7, 5, 4)
    spec, fixed, f, x, rng = _problem(*grid, seed=8, law=law)
    x[:5] = 1e-4  # below the clamp floor
    KE = O.ke_hex8(0.3)
    edof = O.edof_matrix(*grid)
    u = rng.standard_normal(spec.n_dof).astype(udtype).astype(np.float64)
    ref = O.sensitivity(KE, u, edof, x, law, 3.0)
    E = O.stiffness(x, law, 3.0)
    ce_ref = float(np.sum(E * O.element_energy(KE, u, edof)))
    with Problem(spec, fixed, f, device=gpu) as P:
        dc, ce = P.sens(torch.from_numpy(x), torch.from_numpy(u.astype(udtype)), with_energy=True)
        dc = dc.cpu().numpy()
    assert np.abs(dc - ref).max() <= 1e-12 * np.abs(ref).max()
    assert abs(ce - ce_ref) <= 1e-12 * abs(ce_ref)


def test_sens_zero_displacement_gives_zero(gpu):
    """KAT from the reference's kernel test: u = 0 => dc = 0
    (tests/test_aggressive_kernel.py:106-125)."""
    spec, fixed, f, x, rng = _problem(4, 4, 4)
    with Problem(spec, fixed, f, device=gpu) as P:
        dc = P.sens(torch.full((spec.n_elem,), 0.5, dtype=torch.float64),
                    torch.zeros(spec.n_dof, dtype=torch.float32))
    assert float(dc.abs().max()) == 0.0


@pytest.mark.parametrize("grid,rmin", [((9, 7, 5), 1.5), ((6, 6, 6), 2.0), ((5, 3, 8), 1.2),
                                       ((4, 4, 4), 2.6)])
def test_filter_matches_csr(gpu, grid, rmin):
    spec, fixed, f, x, rng = _problem(*grid, seed=9, rmin=rmin)
    H, Hs = O.build_filter(*grid, rmin)
    v = rng.standard_normal(spec.n_elem)
    ref0 = np.asarray(H @ v).ravel() / Hs
    ref1 = O.filter_dc(H, Hs, v)
    with Problem(spec, fixed, f, device=gpu) as P:
        out0 = P.filter(torch.from_numpy(v), 0).cpu().numpy()
        out1 = P.filter(torch.from_numpy(v), 1).cpu().numpy()
    assert np.abs(out0 - ref0).max() <= 1e-15 * np.abs(ref0).max()
    assert np.abs(out1 - ref1).max() <= 1e-15 * np.abs(ref1).max()


@pytest.mark.parametrize("variant", ["cpu", "gpu"])
@pytest.mark.parametrize("with_pres", [False, True])
def test_oc_matches_oracle(gpu, variant, with_pres):
    grid = (10, 6, 5)
    spec, fixed, f, x, rng = _problem(*grid, seed=10, oc=variant, volfrac=0.35)
    H, Hs = O.build_filter(*grid, 1.5)
    dc = -np.abs(rng.standard_normal(spec.n_elem)) * 10 ** rng.uniform(-3, 1, spec.n_elem)
    pres = None
    if with_pres:
        pres = np.zeros(spec.n_elem, bool)
        pres[:17] = True
    oc = O.oc_update_cpu if variant == "cpu" else O.oc_update_gpu
    xn_ref, xp_ref, lmid_ref, steps_ref = oc(x, dc, H, Hs, 0.35, 0.2, pres)
    with Problem(spec, fixed, f, preserved=pres, device=gpu) as P:
        xn, xp, lmid, steps = P.oc(torch.from_numpy(x), torch.from_numpy(dc))
    assert steps == steps_ref
    assert lmid == lmid_ref
    assert np.abs(xn.cpu().numpy() - xn_ref).max() <= 1e-15
    assert np.abs(xp.cpu().numpy() - xp_ref).max() <= 1e-15


def test_compliance_dot(gpu):
    spec, fixed, f, x, rng = _problem(5, 4, 3)
    u = rng.standard_normal(spec.n_dof)
    with Problem(spec, fixed, f, device=gpu) as P:
        c = P.compliance(torch.from_numpy(u))
    assert abs(c - f @ u) <= 1e-14 * np.abs(f).sum() * np.abs(u).max()


@pytest.mark.parametrize("law,oc", [("emin", "cpu"), ("clamp", "gpu")])
def test_simp_one_step_map(gpu, law, oc):
    """One OC iteration from a shared state: density within 1e-5 max-abs."""
    grid = (16, 8, 8)
    spec, fixed, f, x, rng = _problem(*grid, seed=11, law=law, oc=oc, volfrac=0.3)
    prob = O.OracleProblem(*grid, f=f, fixed=fixed, law=law, oc=oc, volfrac=0.3)
    x0 = np.clip(rng.uniform(0.0, 0.6, spec.n_elem), 0, 1)
    H, Hs = O.build_filter(*grid, 1.5)
    xp0 = np.asarray(H @ x0).ravel() / Hs
    xn_ref, xpn_ref, u_ref, c_ref, info = prob.step(x0, xp0, "pcg", 1e-12)
    with Problem(spec, fixed, f, device=gpu) as P:
        r = P.run(1, x0=torch.from_numpy(x0), xphys0=torch.from_numpy(xp0), tol=1e-11,
                  refine=True, inner_tol=1e-4, warm_start=False)
    assert r.status == 0
    assert np.abs(r.x.cpu().numpy() - xn_ref).max() <= 1e-5
    assert np.abs(r.xPhys.cpu().numpy() - xpn_ref).max() <= 1e-5


def test_simp_run_matches_oracle_small(gpu):
    """Free-running 12 OC iterations with FP64-accurate solves: final
    compliance within 1e-6 relative of the FP64 oracle."""
    grid = (16, 8, 8)
    spec, fixed, f, x, rng = _problem(*grid, seed=12, volfrac=0.3)
    prob = O.OracleProblem(*grid, f=f, fixed=fixed, volfrac=0.3)
    xp_ref, c_ref, _ = prob.run(12, "pcg", 1e-12)
    with Problem(spec, fixed, f, device=gpu) as P:
        r = P.run(12, tol=1e-11, refine=True, inner_tol=1e-4)
    assert r.iters_done == 12
    assert abs(r.compliance - c_ref) <= 1e-6 * abs(c_ref)
    assert np.abs(r.xPhys.cpu().numpy() - xp_ref).max() <= 1e-4
