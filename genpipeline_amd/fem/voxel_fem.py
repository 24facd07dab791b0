"""VoxelFEMEvaluator: latent vector -> voxels -> FEM metrics, GPU branch.

Drop-in for genpipeline/fem/voxel_fem.py:403-615 (``evaluate``,
``evaluate_batch``, ``evaluation_history``, ``save_history``). The evaluation
is the reference's GPU branch (:440-500):

  decode z with the caller's VAE -> 2x trilinear down-sampling
  (scipy.ndimage.zoom(0.5, order=1)) -> smooth_voxels (Gaussian sigma 0.5 +
  logistic sharpening 15) -> GPUVoxelFEM.solve ->
  {"stress", "compliance", "mass"} with mass = solid_frac * volume * 7900/1e9
  (volume from ``bbox`` when given, else one mm^3 per voxel).

The CalculiX path (.inp writer, ccx run, .frd parser; :84-399) is not part
of this library. With ``use_gpu=False``, or when the GPU solve raises, the
result is the reference's no-solver penalty: stress and compliance
FEM_SENTINEL, mass 1.0, ``failure_reason`` set (:530-536).
"""
from __future__ import annotations

import concurrent.futures
import json
import logging
from pathlib import Path
from typing import Dict

import numpy as np

from ..pipeline_utils import FEM_SENTINEL, NumpyEncoder, smooth_voxels

logger = logging.getLogger(__name__)


def _mass(voxels: np.ndarray, bbox: dict | None) -> float:
    solid_frac = float((voxels > 0.5).mean())
    D, H, W = voxels.shape
    if bbox:
        vol_mm3 = ((bbox["x"][1] - bbox["x"][0]) * (bbox["y"][1] - bbox["y"][0])
                   * (bbox["z"][1] - bbox["z"][0])) * solid_frac
    else:
        vol_mm3 = D * H * W * solid_frac
    return vol_mm3 * 7900 / 1e9


def decode_to_voxels(z, vae_model) -> np.ndarray:
    """VAE decode, 2x down-sampling and the organic density filter
    (voxel_fem.py:449-462)."""
    import torch
    from scipy.ndimage import zoom

    device = next(vae_model.parameters()).device
    with torch.no_grad():
        zt = torch.from_numpy(np.array(z).reshape(1, -1)).float().to(device)
        voxels = vae_model.decode(zt).cpu().numpy().squeeze()
    return smooth_voxels(zoom(voxels, 0.5, order=1))


class VoxelFEMEvaluator:
    """Evaluator with the reference constructor (voxel_fem.py:406-438).
    ``ccx_cmd`` and ``output_dir`` are accepted and kept for callers that
    read them; nothing is written to ``output_dir``."""

    def __init__(self, ccx_cmd: str = None, output_dir: str = None, fixed_face: str = "x_min",
                 load_face: str = "x_max", force_n: float = 1000.0, vae_model=None,
                 use_gpu: bool = False):
        self.ccx_cmd = ccx_cmd
        self.output_dir = Path(output_dir or "./optimisation_results/voxel_fem")
        self.fixed_face = fixed_face
        self.load_face = load_face
        self.force_n = force_n
        self.vae_model = vae_model
        self.use_gpu = use_gpu
        self.evaluation_history = []
        self._counter = 0

    def evaluate_voxels(self, voxels: np.ndarray, bbox: dict = None) -> Dict[str, float]:
        """FEM metrics of an already decoded and smoothed voxel field."""
        if not self.use_gpu:
            self._counter += 1
            return {"stress": FEM_SENTINEL, "compliance": FEM_SENTINEL, "mass": 1.0,
                    "failure_reason": "no_ccx"}
        from .gpu_fem_solver import GPUVoxelFEM
        try:
            fem = GPUVoxelFEM(voxels=voxels, fixed_face=self.fixed_face, load_face=self.load_face,
                              force_n=self.force_n).solve()
        except Exception as e:  # the reference falls back to ccx here; there is none
            logger.warning(f"[VoxelFEM] GPU FEM failed: {e}")
            self._counter += 1
            return {"stress": FEM_SENTINEL, "compliance": FEM_SENTINEL, "mass": 1.0,
                    "failure_reason": "gpu_fem_error"}
        result = {"stress": fem["stress_max"], "compliance": fem["compliance"],
                  "mass": _mass(voxels, bbox)}
        self._counter += 1
        self.evaluation_history.append(result)
        return result

    def evaluate(self, z: np.ndarray, vae_model, bbox: dict = None) -> Dict[str, float]:
        """Decode z -> voxels -> FEM -> {stress, compliance, mass}."""
        return self.evaluate_voxels(decode_to_voxels(z, vae_model), bbox)

    def evaluate_batch(self, param_list: list) -> list:
        """One FEMResult per entry of ``param_list`` (reads ``latent_z``);
        voxel_fem.py:559-611."""
        if self.vae_model is None:
            raise RuntimeError("VoxelFEMEvaluator.evaluate_batch requires vae_model set at construction")
        from ..schema import FEMResult, OptimizationSample

        n_workers = getattr(self, "n_workers", 4)
        results = [None] * len(param_list)

        def eval_one(idx, p):
            z = getattr(p, "latent_z", None)
            if z is None:
                return idx, FEMResult(stress_max=FEM_SENTINEL, compliance=FEM_SENTINEL, mass=1.0,
                                      success=False), None
            r = self.evaluate(np.asarray(z), self.vae_model)
            return idx, FEMResult(stress_max=float(r.get("stress", FEM_SENTINEL)),
                                  compliance=float(r.get("compliance", FEM_SENTINEL)),
                                  mass=float(r.get("mass", 1.0)),
                                  success=r.get("failure_reason") is None), z

        with concurrent.futures.ThreadPoolExecutor(max_workers=n_workers) as ex:
            futs = [ex.submit(eval_one, i, p) for i, p in enumerate(param_list)]
            for fut in concurrent.futures.as_completed(futs):
                idx, res, z = fut.result()
                results[idx] = res
                if z is not None:
                    self.evaluation_history.append(OptimizationSample(
                        parameters=param_list[idx], result=res,
                        latent_z=z.tolist() if hasattr(z, "tolist") else list(z)))
        return results

    def save_history(self, path: str):
        with open(path, "w") as f:
            json.dump(self.evaluation_history, f, indent=2, cls=NumpyEncoder)
