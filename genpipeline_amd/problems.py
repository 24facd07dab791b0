"""Seeded synthetic SIMP workloads.

Every generator draws from ``numpy.random.default_rng(seed)`` in a fixed
order, so a (generator, seed) pair names one exact problem.

* ``cantilever_32x16x16``  32x16x16 cantilever, clamp i=0, unit -z load at
  node (32, 8, 0) (DOF 28154), volfrac 0.3.
* ``point_loads``          clamp i=0; n seeded unit point loads on face
  i=nx: nodes = rng.choice((ny+1)(nz+1), n, replace=False) in (j,k)
  row-major order, directions g/|g| with g = rng.standard_normal((n,3)).
* ``traction``             total magnitude 1 in -z on face i=nx, bilinear
  nodal weights (1/2 at face edges, 1 inside), normalised to sum 1.
* ``smoothed_density``     Gaussian-smoothed seeded noise (sigma voxels),
  clipped to [0.01,1] and rescaled to a target mean.
* ``batch_specs``          the uneven 48-problem batch: size class from
  {32^3, 48x24x24, 64x32x32, 64^3}, volfrac U[0.2,0.4], 1-4 point loads.
* ``serving_requests``     Poisson arrivals of FEM/SIMP evaluation requests.
* ``SyntheticDecoder``     stand-in for the design VAE's decoder: seeded,
  smooth random solids of the decoder's output shape.
* ``bo_latents``           a batch of seeded latent vectors (BO candidates).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .bcs import clamp_x0
from .problem import ProblemSpec

SIZE_CLASSES = [(32, 32, 32), (48, 24, 24), (64, 32, 32), (64, 64, 64)]


def node_id(i, j, k, ny, nz):
    return i * (ny + 1) * (nz + 1) + j * (nz + 1) + k


@dataclass
class Workload:
    spec: ProblemSpec
    fixed: np.ndarray
    f: np.ndarray
    x0: np.ndarray | None = None
    n_iters: int = 50
    tol: float = 1e-6
    name: str = ""


def cantilever_32x16x16(law: str = "emin") -> Workload:
    spec = ProblemSpec(32, 16, 16, volfrac=0.3, law=law)
    f = np.zeros(spec.n_dof)
    f[3 * node_id(32, 8, 0, 16, 16) + 2] = -1.0
    return Workload(spec, clamp_x0(32, 16, 16), f, n_iters=50, tol=1e-8, name="cantilever_32x16x16")


def point_load_vector(nx, ny, nz, rng, n_loads):
    f = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    picks = rng.choice((ny + 1) * (nz + 1), n_loads, replace=False)
    g = rng.standard_normal((n_loads, 3))
    dirs = g / np.linalg.norm(g, axis=1, keepdims=True)
    for p, d in zip(picks, dirs):
        j, k = divmod(int(p), nz + 1)
        n = node_id(nx, j, k, ny, nz)
        f[3 * n:3 * n + 3] += d
    return f


def point_loads(nx=64, ny=64, nz=64, seed=0, n_loads=4, volfrac=0.3, n_iters=100,
                tol=1e-6) -> Workload:
    rng = np.random.default_rng(seed)
    spec = ProblemSpec(nx, ny, nz, volfrac=volfrac)
    f = point_load_vector(nx, ny, nz, rng, n_loads)
    return Workload(spec, clamp_x0(nx, ny, nz), f, n_iters=n_iters, tol=tol,
                    name=f"point_loads_{nx}x{ny}x{nz}_s{seed}")


def traction(nx, ny, nz, total=1.0) -> np.ndarray:
    wj = np.ones(ny + 1)
    wj[[0, -1]] = 0.5
    wk = np.ones(nz + 1)
    wk[[0, -1]] = 0.5
    w = np.outer(wj, wk).ravel()
    w = w / w.sum() * total
    f = np.zeros(3 * (nx + 1) * (ny + 1) * (nz + 1))
    base = node_id(nx, 0, 0, ny, nz)
    f[3 * (base + np.arange(w.size)) + 2] = -w
    return f


def smoothed_density(shape, seed, sigma=2.0, mean=0.25, lo=0.01, hi=1.0) -> np.ndarray:
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    g = gaussian_filter(rng.standard_normal(shape), sigma, mode="reflect")
    g = (g - g.mean()) / g.std()
    x = np.clip(mean + 0.15 * g, lo, hi)
    for _ in range(200):
        if abs(x.mean() - mean) < 1e-12:
            break
        x = np.clip(x * mean / x.mean(), lo, hi)
    return x.ravel()


def mixed_128x64x64(seed=0, n_iters=50) -> Workload:
    spec = ProblemSpec(128, 64, 64, volfrac=0.25)
    x0 = smoothed_density((128, 64, 64), seed)
    return Workload(spec, clamp_x0(128, 64, 64), traction(128, 64, 64), x0=x0, n_iters=n_iters,
                    tol=1e-6, name=f"mixed_128x64x64_s{seed}")


def batch_specs(seed=0, n=48, n_iters=60, tol=1e-6) -> list[Workload]:
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n):
        cls = int(rng.integers(4))
        vf = float(rng.uniform(0.2, 0.4))
        nl = int(rng.integers(1, 5))
        nx, ny, nz = SIZE_CLASSES[cls]
        f = point_load_vector(nx, ny, nz, rng, nl)
        out.append(Workload(ProblemSpec(nx, ny, nz, volfrac=vf), clamp_x0(nx, ny, nz), f,
                            n_iters=n_iters, tol=tol, name=f"batch{b}_{nx}x{ny}x{nz}"))
    return out


@dataclass
class Request:
    rid: int
    t_arrive: float
    shape: tuple
    density: np.ndarray  # or a device tensor of the grid's elements
    kind: str           # "solve" | "simp"
    volfrac: float
    rhs: object = None   # "solve" only: device load vector replacing the problem's
    tol: float = None    # "solve" only: PCG tolerance replacing the server's


def serving_requests(lam: float, n: int, seed=0, p_small=0.6, p_solve=0.5,
                     shapes=((32, 32, 32), (64, 64, 64))) -> list[Request]:
    """Open-loop Poisson arrivals (rate lam per second); the grid is shapes[0]
    with probability p_small, else shapes[1]; densities are smoothed noise
    (sigma 2) thresholded at the (1-phi) quantile with solid fraction
    phi ~ U[0.15,0.45]; the kind is "solve" with probability p_solve."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for r in range(n):
        t += float(rng.exponential(1.0 / lam))
        shape = tuple(shapes[0]) if rng.uniform() < p_small else tuple(shapes[1])
        phi = float(rng.uniform(0.15, 0.45))
        g = gaussian_filter(rng.standard_normal(shape), 2.0, mode="reflect")
        thr = np.quantile(g, 1.0 - phi)
        dens = (g > thr).astype(np.float64).ravel()
        kind = "solve" if rng.uniform() < p_solve else "simp"
        out.append(Request(r, t, shape, dens, kind, phi))
    return out


class SyntheticDecoder:
    """Stands in for DesignVAE.decode (genpipeline/vae_design_model.py:190-212:
    sigmoid of the decoder logits, (B, 1, res, res, res) float32) where no
    trained checkpoint exists. decode(z) = sigmoid(gain (h + bias)) with h a
    fixed seeded linear map of z onto a coarse^3 grid, trilinearly up-sampled
    to res^3: smooth random solids, about half solid at bias 0."""

    def __init__(self, latent_dim: int = 32, res: int = 64, coarse: int = 8, seed: int = 0,
                 gain: float = 6.0, bias: float = 0.0, device=None):
        import torch
        g = torch.Generator().manual_seed(int(seed))
        w = torch.randn(latent_dim, coarse ** 3, generator=g) / float(latent_dim) ** 0.5
        self._w = torch.nn.Parameter(w.to(device) if device is not None else w,
                                     requires_grad=False)
        self.latent_dim, self.res, self.coarse = latent_dim, res, coarse
        self.gain, self.bias = float(gain), float(bias)

    def parameters(self):
        return iter([self._w])

    def decode(self, z):
        import torch
        c = self.coarse
        h = (z.float().to(self._w.device) @ self._w).reshape(-1, 1, c, c, c)
        h = torch.nn.functional.interpolate(h, size=(self.res,) * 3, mode="trilinear",
                                            align_corners=True)
        return torch.sigmoid(self.gain * (h + self.bias))


def bo_latents(n: int, latent_dim: int = 32, seed: int = 0) -> np.ndarray:
    """n latent vectors drawn N(0, 1): rng.standard_normal((n, latent_dim))."""
    return np.random.default_rng(seed).standard_normal((n, latent_dim))
