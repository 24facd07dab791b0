"""CPU oracle of the FEM evaluator's voxel pre-processing -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` may import this module; the product never does.

``evaluator_voxels`` is the reference chain itself (genpipeline/fem/
voxel_fem.py:452-462 and smooth_voxels, genpipeline/pipeline_utils.py:55-76):
``scipy.ndimage.zoom(v, 0.5, order=1)``, ``gaussian_filter(v, 0.5)`` and the
float32 logistic, computed with the same third-party routines the reference
calls. The reference pins no SciPy version (requirements.txt:12 says
``scipy``); this image has SciPy 1.15.3.

``zoom1_restated`` and ``gaussian_restated`` restate SciPy's published
algorithms for those two calls (ndimage/src/ni_interpolation.c NI_ZoomShift
with order 1; ni_filters.c NI_Correlate1D, symmetric-kernel branch): the
coordinate map o * (n-1)/(m-1), the 'constant'-mode rule that a coordinate
rounded past n-1 takes cval, the footprint and weight order of the
interpolation sum, the pairwise (v[-j] + v[j]) w[j] order of the filter sum,
'reflect' boundaries, and float32 rounding after the zoom and after each
filter axis. The HIP kernels follow these steps.

How each function is pinned (tests/test_oracle.py):
  zoom1_restated     bit-equal to scipy.ndimage.zoom over random shapes,
                     including 32 -> 16, where 15 * (31/15) rounds past 31 and
                     SciPy writes cval; an affine field is reproduced at the
                     mapped coordinates (closed form).
  gaussian_restated  bit-equal to scipy.ndimage.gaussian_filter; a constant
                     field stays constant; the impulse response of an
                     interior voxel is the outer product of the taps.
  evaluator_voxels   the reference's own calls; logistic(0.5) = 0.5; the
                     two exp variants agree within one float32 ulp.
"""
from __future__ import annotations

import numpy as np


def evaluator_voxels(decoded: np.ndarray, factor: float = 0.5, sigma: float = 0.5,
                     sharpness: float = 15.0, exp: str = "numpy") -> np.ndarray:
    """The reference pre-processing of one decoded (D, H, W) float32 field
    (voxel_fem.py:458-462, pipeline_utils.py:71-75).

    exp="numpy": the logistic exactly as the reference evaluates it, with
    NumPy's float32 exp (a SIMD routine chosen by the host CPU; it differs
    from the correctly rounded value in the last bit for about 40 % of the
    arguments here). exp="double": the same float32 steps with exp evaluated
    in float64 and rounded once, as the HIP kernel evaluates it."""
    from scipy.ndimage import gaussian_filter, zoom
    v = zoom(np.asarray(decoded, np.float32), factor, order=1)
    v = gaussian_filter(v, sigma=sigma)
    if exp == "numpy":
        return 1.0 / (1.0 + np.exp(-sharpness * (v - 0.5)))
    a = np.float32(-sharpness) * (v - np.float32(0.5))
    e = np.exp(a.astype(np.float64)).astype(np.float32)
    return np.float32(1.0) / (np.float32(1.0) + e)


def taps(sigma: float = 0.5, truncate: float = 4.0) -> np.ndarray:
    """gaussian_filter1d's weights: radius int(truncate sigma + 0.5),
    exp(-x^2 / (2 sigma^2)) normalised to sum 1."""
    r = int(truncate * float(sigma) + 0.5)
    x = np.arange(-r, r + 1)
    phi = np.exp(-0.5 / (sigma * sigma) * x ** 2)
    return phi / phi.sum()


def zoom1_restated(a: np.ndarray, out_shape) -> np.ndarray:
    """Order-1 zoom of a 3-D float array to ``out_shape`` (grid_mode False,
    mode 'constant', cval 0), in SciPy's operation order."""
    a64 = np.asarray(a).astype(np.float64)
    idx, wts, outside = [], [], []
    for n, m in zip(a.shape, out_shape):
        z = (n - 1) / (m - 1) if m > 1 else 1.0
        c = np.arange(m, dtype=np.float64) * z
        s = np.floor(c).astype(np.int64)
        t = c - s
        idx.append((s, np.minimum(s + 1, n - 1)))
        wts.append((1.0 - t, t))
        outside.append(c > n - 1)
    out = np.zeros(tuple(out_shape), np.float64)
    for p in range(2):
        for q in range(2):
            for r in range(2):  # footprint order: last axis fastest
                v = a64[np.ix_(idx[0][p], idx[1][q], idx[2][r])]
                term = v * wts[0][p][:, None, None]
                term = term * wts[1][q][None, :, None]
                term = term * wts[2][r][None, None, :]
                out = out + term
    for ax in range(3):
        sl = [slice(None)] * 3
        sl[ax] = outside[ax]
        out[tuple(sl)] = 0.0
    return out.astype(np.asarray(a).dtype)


def _reflect(i: np.ndarray, n: int) -> np.ndarray:
    if n == 1:
        return np.zeros_like(i)
    i = np.mod(i, 2 * n)
    return np.where(i >= n, 2 * n - 1 - i, i)


def gaussian_restated(a: np.ndarray, sigma: float = 0.5) -> np.ndarray:
    """gaussian_filter(a, sigma) on a 3-D float32 array: one correlation per
    axis (0, 1, 2), each summed in double as v[0] w[0] + sum_{j=-r..-1}
    (v[j] + v[-j]) w[j] and stored as float32, 'reflect' boundaries."""
    w = taps(sigma)
    r = (len(w) - 1) // 2
    v = np.asarray(a)
    for ax in range(3):
        n = v.shape[ax]
        base = np.arange(n)
        vd = v.astype(np.float64)
        acc = np.take(vd, base, axis=ax) * w[r]
        for j in range(-r, 0):
            lo = np.take(vd, _reflect(base + j, n), axis=ax)
            hi = np.take(vd, _reflect(base - j, n), axis=ax)
            acc = acc + (lo + hi) * w[r + j]
        v = acc.astype(np.asarray(a).dtype)
    return v
