"""Why the 16-bit PCG variant stores only x and p in 16 bits: an FP64
emulation of the single-kernel PCG (Chronopoulos-Gear recurrence) with its
stored vectors rounded to fp16, inside FP64 iterative refinement (outer
1e-9, inner 1e-3), on the 128x64x64 configuration's field family
(smoothed noise floored at 0.01, modified SIMP, traction on x = max) at
32x16x16.

The K.u operand of the fused iteration is z = M r, formed in registers from
the recurrence vectors r, w, s; the element matrix is eight modal scalars in
registers. Neither is a stored array, so "16-bit K.u operand" means 16-bit
r, w, s. Measured here: storing x and p in fp16 converges to the FP64
answer (more inner iterations than FP32); storing r, w and s in fp16 breaks
the recurrence down during the first inner solve, with the inner
right-hand side scaled to either |f| or RMS 1.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import simp_oracle as O


def _system():
    from genpipeline_amd.bcs import clamp_x0
    from genpipeline_amd.problems import smoothed_density, traction
    nx, ny, nz = 32, 16, 16
    x = smoothed_density((nx, ny, nz), 0)
    assert x.min() >= 0.01 - 1e-15
    f, fixed = traction(nx, ny, nz), clamp_x0(nx, ny, nz)
    K = O.assemble_K(O.ke_hex8(0.3), O.stiffness(x, "emin", 3.0), O.edof_matrix(nx, ny, nz),
                     f.size)
    free = np.ones(f.size, bool)
    free[fixed] = False
    return K[free][:, free].tocsr(), f[free]


def _q(v, dt):
    return v if dt is None else v.astype(dt).astype(np.float64)


def _inner(A, Minv, rhs, tol, rdt, xdt, sig, maxit=20000):
    """Chronopoulos-Gear PCG; r, w, s stored as rdt, x, p as xdt. Returns
    (correction, iterations) with iterations < 0 on breakdown."""
    r = _q(rhs * sig, rdt)
    z = Minv * r
    w = _q(A @ z, rdt)
    g, d = r @ z, z @ w
    a = g / d
    p, s, x = _q(z, xdt), w.copy(), np.zeros_like(r)
    bn = np.linalg.norm(r)
    for it in range(1, maxit + 1):
        with np.errstate(all="ignore"):
            x = _q(x + a * p, xdt)
            r = _q(r - a * s, rdt)
            if np.linalg.norm(r) <= tol * bn:
                return x / sig, it
            z = Minv * r
            w = _q(A @ z, rdt)
            gn, d = r @ z, z @ w
            bta = gn / g
            den = d - bta * gn / a
            if not den > 0:
                return x / sig, -it
            g, a = gn, gn / den
            p, s = _q(z + bta * p, xdt), _q(w + bta * s, rdt)
    return x / sig, maxit


def _refine(A, b, rdt, xdt, scale):
    Minv = 1.0 / A.diagonal()
    u, its, bb = np.zeros_like(b), 0, np.linalg.norm(b)
    for k in range(60):
        r = b - A @ u
        if np.linalg.norm(r) <= 1e-9 * bb:
            return u, its, k
        sig = (bb if scale == "norm" else np.sqrt(r.size)) / np.linalg.norm(r)
        d, it = _inner(A, Minv, r, 1e-3, rdt, xdt, sig)
        if it < 0:
            return None, its - it, k
        its += it
        u = u + d
    return u, its, 60


def test_16bit_storage_of_the_recurrence_breaks_down():
    A, b = _system()
    ud = spla.spsolve(A.tocsc(), b)
    u32, it32, _ = _refine(A, b, np.float32, np.float32, "norm")
    uxp, itxp, _ = _refine(A, b, None, np.float16, "norm")
    for u in (u32, uxp):
        assert u is not None and np.abs(u - ud).max() <= 1e-10 * np.abs(ud).max()
    assert itxp > it32
    for scale in ("norm", "rms"):
        u, its, outer = _refine(A, b, np.float16, np.float16, scale)
        assert u is None and outer == 0, (scale, its, outer)  # breakdown in the first inner solve
    u, its, outer = _refine(A, b, np.float16, None, "rms")
    assert u is None and outer == 0
