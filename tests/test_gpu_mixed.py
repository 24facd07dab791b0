"""16-bit PCG storage inside FP64 iterative refinement: the refined answer
is the FP64 answer (against the oracle's direct solve), and a SIMP run with
it matches the FP64 oracle to the same bars as the FP32 path."""
import numpy as np
import pytest
import torch

from genpipeline_amd import Problem, ProblemSpec
from genpipeline_amd.bcs import clamp_x0
from genpipeline_amd.problems import smoothed_density, traction
from oracle import simp_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,lo", [("bf16", 0.01), ("fp16", 0.3)])
def test_refined_16bit_solve_matches_direct(gpu, prec, lo):
    """bf16 keeps float's exponent range (densities down to 0.01, E ~ 1e-6);
    fp16 is exercised on a field whose solution fits its range."""
    nx, ny, nz = 16, 8, 8
    spec = ProblemSpec(nx, ny, nz, volfrac=0.25)
    f, fixed = traction(nx, ny, nz), clamp_x0(nx, ny, nz)
    x = smoothed_density((nx, ny, nz), 4, lo=lo, mean=max(0.25, lo + 0.1))
    P = Problem(spec, fixed, f, device=gpu)
    E = P.stiffness(x)
    u, st = P.pcg(E, tol=1e-10, precision=prec, refine=True, inner_tol=1e-2, max_outer=60)
    assert st.status == 0 and st.relres <= 1e-10 and st.outer >= 2
    Po = O.OracleProblem(nx, ny, nz, f, fixed)
    Ee = E.double().cpu().numpy()
    K = O.assemble_K(O.ke_hex8(0.3), Ee, O.edof_matrix(nx, ny, nz), spec.n_dof)
    ud, _ = O.solve(K, f, fixed, "direct")
    assert u.dtype == torch.float64
    assert np.abs(u.cpu().numpy() - ud).max() <= 1e-8 * np.abs(ud).max()
    P.close()


def test_simp_run_bf16_matches_oracle(gpu):
    nx, ny, nz = 12, 6, 6
    spec = ProblemSpec(nx, ny, nz, volfrac=0.3)
    f = np.zeros(spec.n_dof)
    f[3 * (spec.n_nodes - 1 - (nz + 1) * (ny // 2)) + 2] = -1.0
    fixed = clamp_x0(nx, ny, nz)
    P = Problem(spec, fixed, f, device=gpu)
    r = P.run(4, tol=1e-10, precision="bf16", refine=True, inner_tol=1e-2, max_outer=60)
    ref = O.OracleProblem(nx, ny, nz, f=f, fixed=fixed, volfrac=0.3)
    xp, c, _ = ref.run(4, "direct")
    assert r.status == 0
    assert abs(r.compliance - c) <= 1e-6 * abs(c)
    assert np.abs(r.xPhys.cpu().numpy() - xp).max() <= 1e-5
    P.close()


def test_16bit_needs_refinement(gpu):
    from genpipeline_amd import _lib
    nx, ny, nz = 6, 4, 4
    spec = ProblemSpec(nx, ny, nz)
    P = Problem(spec, clamp_x0(nx, ny, nz), traction(nx, ny, nz), device=gpu)
    E = P.stiffness(np.full(spec.n_elem, 0.5))
    with pytest.raises(_lib.GpError, match="refinement"):
        P.pcg(E, precision="bf16", refine=False)
    P.close()


@pytest.mark.parametrize("K", [3, 4, 7])
@pytest.mark.parametrize("prec,bound", [("fp16", 3e-3), ("bf16", 2e-2)])
def test_capped_16bit_inner_solve_tracks_fp32(gpu, prec, bound, K):
    """One inner solve capped at K iterations (max_outer 1, u = its
    correction). x and p are updated two iterations at a time, and after an
    odd K the last update is completed once the solve ends. The 16-bit
    iterate equals the float one up to 16-bit rounding, at odd and even K. The
    float iterate of one iteration fewer differs from it by far more, so a
    missing last update would fail the bound."""
    nx, ny, nz = 16, 8, 8
    P = Problem(ProblemSpec(nx, ny, nz), clamp_x0(nx, ny, nz), traction(nx, ny, nz), device=gpu)
    E = P.stiffness(torch.full((P.spec.n_elem,), 0.5, dtype=torch.float64, device=gpu))

    def solve(p, k):
        u, st = P.pcg(E, tol=1e-14, max_iter=k, precision=p, refine=True, inner_tol=1e-14,
                      max_outer=1)
        assert st.iters == k and st.outer == 1, (p, k, st)
        return u

    ref = solve("fp32", K)
    scale = float(ref.abs().max())
    d_lp = float((solve(prec, K) - ref).abs().max()) / scale
    d_prev = float((solve("fp32", K - 1) - ref).abs().max()) / scale
    print(f"[capped {prec} K={K}] 16-bit vs float {d_lp:.2e}, one iteration fewer {d_prev:.2e}")
    assert d_lp <= bound, d_lp
    assert d_prev >= 3 * bound, d_prev
    P.close()
