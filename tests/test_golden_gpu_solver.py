"""The oracle's and the product's restatements of the reference GPU solver's
host-side functions against golden vectors of the reference itself
(tests/golden/make_golden_gpu.py: SIMPSolverGPU imported on CPU torch).

Pinned here: ``_get_bcs`` (simp_solver_gpu.py:138-311), ``_build_filter``
(:869-930), ``_filter_dc`` (:796-799), ``_oc_update`` (:801-841) and the
``run`` driver (:71-136, with the reference's ``_solve`` replaced by a direct
solve because its CPU branch reads an undefined name at :677).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import simp_oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden" / "ref_simp_gpu_golden.npz"


@pytest.fixture(scope="module")
def gold():
    meta = json.loads(GOLDEN.with_suffix(".json").read_text())
    with np.load(GOLDEN, allow_pickle=False) as d:
        arrays = {k: np.array(d[k]) for k in d.files}
    return arrays, meta


def _grid(tag):
    return tuple(int(v) for v in tag.split("/")[2].split("x"))


def _bc_impls():
    from genpipeline_amd.bcs import simp_gpu_bcs
    return [("oracle", O.get_bcs_gpu), ("product", simp_gpu_bcs)]


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_gpu_bcs_match_golden_f64(gold, impl):
    """float64: load values and fixed DOFs bit for bit, fixed DOFs in the
    reference's order (x, then y, then z components of the face nodes)."""
    fn = dict(_bc_impls())[impl]
    arrays, meta = gold
    n = 0
    for tag, bc in meta["bcs"].items():
        if not tag.endswith("/f64"):
            continue
        nx, ny, nz = _grid(tag)
        f, fixed = fn(bc, nx, ny, nz, 1.5)
        nzi = np.nonzero(f)[0]
        assert np.array_equal(nzi, arrays[tag + "/f_idx"]), tag
        assert np.array_equal(f[nzi], arrays[tag + "/f_val"]), tag
        assert np.array_equal(np.asarray(fixed, np.int64), arrays[tag + "/fixed"]), tag
        n += 1
    assert n >= 60


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_gpu_bcs_match_golden_f32(gold, impl):
    """The reference's default float32 dtype: the float64 loads rounded once
    to float32 are the reference's float32 loads (equal, or within one
    float32 ulp where the reference subtracts several shares in float32)."""
    fn = dict(_bc_impls())[impl]
    arrays, meta = gold
    for tag, bc in meta["bcs"].items():
        if not tag.endswith("/f32"):
            continue
        nx, ny, nz = _grid(tag)
        f, fixed = fn(bc, nx, ny, nz, 1.5)
        f32 = f.astype(np.float32)
        nzi = np.nonzero(f32)[0]
        assert np.array_equal(nzi, arrays[tag + "/f_idx"]), tag
        g = arrays[tag + "/f_val"].astype(np.float32)
        ulp = np.spacing(np.abs(g))
        assert np.all(np.abs(f32[nzi] - g) <= ulp), tag
        assert np.array_equal(np.asarray(fixed, np.int64), arrays[tag + "/fixed"]), tag


@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_gpu_bcs_reject_what_the_reference_rejects(gold, impl):
    """A list load_dof reaches `tensor + list` in the reference: TypeError."""
    fn = dict(_bc_impls())[impl]
    from tests.golden.make_golden_gpu import BCS
    _, meta = gold
    assert meta["bcs_errors"]
    for tag, err in meta["bcs_errors"].items():
        assert err == "TypeError", tag
        key = tag.split("/")[1]
        nx, ny, nz = _grid(tag)
        with pytest.raises(TypeError):
            fn(dict(BCS[key]), nx, ny, nz, 1.5)


def test_filter_matches_golden(gold):
    """H (weights rmin - |d| over the ceil(rmin) box), Hs and the adjoint
    filter H^T (dc / Hs), for rmin 1.5, 2.0 and 2.5."""
    arrays, meta = gold
    for tag, c in meta["filter"].items():
        H, Hs = O.build_filter(c["nx"], c["ny"], c["nz"], c["rmin"], "f32")
        H64, _ = O.build_filter(c["nx"], c["ny"], c["nz"], c["rmin"])
        # float32-rounded weights even in the float64 solver (1.5 - sqrt(2) differs)
        assert not np.array_equal(H.toarray(), H64.toarray())
        idx, val = arrays[tag + "/H_idx"], arrays[tag + "/H_val"]
        keep = val != 0.0  # the reference keeps explicit zeros at |d| = rmin
        n = H.shape[0]
        G = np.zeros((n, n))
        G[idx[0][keep], idx[1][keep]] = val[keep]
        assert np.array_equal(H.toarray(), G), tag
        assert np.allclose(Hs, arrays[tag + "/Hs"], rtol=2e-16 * 20, atol=0), tag
        assert np.allclose(Hs.astype(np.float32), arrays[tag + "/Hs_f32"], rtol=1e-6, atol=0)
        dcf = O.filter_dc(H, Hs, arrays[tag + "/dc_in"])
        assert np.allclose(dcf, arrays[tag + "/dc_out"], rtol=1e-14, atol=1e-15), tag


def test_oc_update_gpu_matches_golden(gold):
    """GPU OC constants (l2 = 2 max(max(-dc), 1), <= 30 steps, relative gap
    1e-5, floor 1e-15), with and without preserved elements."""
    arrays, meta = gold
    for tag, c in meta["oc"].items():
        H, Hs = O.build_filter(c["nx"], c["ny"], c["nz"], c["rmin"], "f32")
        pres = None
        if c["preserved"]:
            m = np.zeros((c["nx"], c["ny"], c["nz"]), bool)
            m[c["nx"] - 2:, :, :2] = True
            pres = m.ravel()
        xnew, xp, _, _ = O.oc_update_gpu(arrays[tag + "/x"], arrays[tag + "/dc"], H, Hs,
                                         c["volfrac"], 0.2, pres)
        assert np.abs(xnew - arrays[tag + "/xnew"]).max() <= 1e-14, tag
        assert np.abs(xp - arrays[tag + "/xphys"]).max() <= 1e-14, tag


def _golden_problem(r):
    pres = None
    if r["preserved"]:
        pres = np.zeros((r["nx"], r["ny"], r["nz"]), bool)
        pres[r["nx"] - 2:, :, :2] = True
    return O.reference_problem(r["nx"], r["ny"], r["nz"], bcs=r["bcs"], force_mag=r["force"],
                               volfrac=r["volfrac"], preserved=pres, ke=r["table"], law="clamp",
                               oc="gpu", gpu_bcs=True)


def test_simp_gpu_run_matches_golden(gold):
    """The driver: x_init rescale and clip, warm starts, compliance solve
    after every update, early stop (the 40-iteration default run stops at
    28), reference table with the reference's assembly orientation."""
    arrays, meta = gold
    assert meta["runs"]["default_10x5x4"]["iters_done"] < meta["runs"]["default_10x5x4"]["n_iters"]
    for name, r in meta["runs"].items():
        P = _golden_problem(r)
        xi = arrays.get(f"run/{name}/x_init")
        xp, comp, done, hist = O.simp_gpu_run(P, r["n_iters"], r["volfrac"], x_init=xi,
                                              method="direct")
        assert done == r["iters_done"], name
        gh = arrays[f"run/{name}/compliance_history"]
        assert np.allclose(hist, gh, rtol=1e-9, atol=0), name
        assert comp == pytest.approx(float(arrays[f"run/{name}/compliance"]), rel=1e-9), name
        assert np.abs(xp - arrays[f"run/{name}/xPhys"].ravel()).max() < 1e-9, name


def test_reference_table_needs_its_assembly_orientation(gold):
    """With the asymmetric reference table the two solvers assemble
    different matrices: the CPU solver's KE^T orientation does not reproduce
    the GPU solver's run."""
    arrays, meta = gold
    r = meta["runs"]["default_8x4x4_reftable"]
    P = _golden_problem(r)
    P.assembly = "cpu"
    _, comp, _, _ = O.simp_gpu_run(P, r["n_iters"], r["volfrac"], method="direct")
    assert comp != pytest.approx(float(arrays["run/default_8x4x4_reftable/compliance"]), rel=1e-3)
