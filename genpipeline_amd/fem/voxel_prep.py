"""Decoded fields -> FEM voxel densities on the device (``gp_voxel_prep``).

The evaluator's pre-processing (genpipeline/fem/voxel_fem.py:452-462 and
smooth_voxels, genpipeline/pipeline_utils.py:55-76) is

    voxels = scipy.ndimage.zoom(decoded, 0.5, order=1)
    voxels = scipy.ndimage.gaussian_filter(voxels, sigma=0.5)
    voxels = 1 / (1 + exp(-15 (voxels - 0.5)))

on float32 arrays. ``prep_voxels`` runs that chain for a whole batch of
decoded fields with the HIP kernels of ``gp_voxel.hip``, in scipy's order of
operations and float32 rounding points. The output shape and the Gaussian
taps are computed here exactly as scipy computes them.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib as L


def zoom_shape(shape, factor: float = 0.5) -> tuple:
    """Output shape of scipy.ndimage.zoom: round(n * factor) per axis, with
    Python's round (halves to even)."""
    return tuple(int(round(n * factor)) for n in shape)


def gaussian_taps(sigma: float = 0.5, truncate: float = 4.0) -> np.ndarray:
    """The 2r+1 weights gaussian_filter correlates with on every axis:
    r = int(truncate * sigma + 0.5), w_x = exp(-x^2 / (2 sigma^2)) / sum."""
    sigma = float(sigma)
    r = int(truncate * sigma + 0.5)
    x = np.arange(-r, r + 1)
    phi = np.exp(-0.5 / (sigma * sigma) * x ** 2)
    return phi / phi.sum()


def prep_voxels(decoded: torch.Tensor, factor: float = 0.5, sigma: float = 0.5,
                sharpness: float = 15.0) -> torch.Tensor:
    """(B, D, H, W) or (B, 1, D, H, W) float32 device tensor -> (B, od, oh, ow)
    float32 voxel densities on the same device (od = round(D * factor))."""
    if not isinstance(decoded, torch.Tensor) or not decoded.is_cuda:
        raise RuntimeError("prep_voxels needs the decoded fields as a ROCm device tensor")
    x = decoded
    if x.dim() == 5:
        if x.shape[1] != 1:
            raise ValueError(f"expected one channel, got shape {tuple(x.shape)}")
        x = x[:, 0]
    if x.dim() != 4:
        raise ValueError(f"expected (B, D, H, W) fields, got shape {tuple(decoded.shape)}")
    x = x.to(torch.float32).contiguous()
    B, D, H, W = x.shape
    od, oh, ow = zoom_shape((D, H, W), factor)
    if min(od, oh, ow) < 1:
        raise ValueError(f"zoom {factor} of {(D, H, W)} leaves an empty axis")
    taps = gaussian_taps(sigma)
    r = (len(taps) - 1) // 2
    out = torch.empty((B, od, oh, ow), dtype=torch.float32, device=x.device)
    work = torch.empty_like(out)
    with torch.cuda.device(x.device):
        L.call("gp_voxel_prep", C.c_void_p(x.data_ptr()), B, D, H, W, od, oh, ow,
               taps.ctypes.data_as(C.c_void_p), r, C.c_double(sharpness),
               C.c_void_p(work.data_ptr()), C.c_void_p(out.data_ptr()),
               C.c_void_p(torch.cuda.current_stream(x.device).cuda_stream))
    return out
