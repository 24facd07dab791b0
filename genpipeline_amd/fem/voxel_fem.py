"""VoxelFEMEvaluator: latent vector -> voxels -> FEM metrics, GPU branch.

Drop-in for genpipeline/fem/voxel_fem.py:403-615 (``evaluate``,
``evaluate_batch``, ``evaluation_history``, ``save_history``). The evaluation
is the reference's GPU branch (:440-500):

  decode z with the caller's VAE -> 2x trilinear down-sampling
  (scipy.ndimage.zoom(0.5, order=1)) -> smooth_voxels (Gaussian sigma 0.5 +
  logistic sharpening 15) -> GPUVoxelFEM.solve ->
  {"stress", "compliance", "mass"} with mass = solid_frac * volume * 7900/1e9
  (volume from ``bbox`` when given, else one mm^3 per voxel).

Down-sampling and smoothing run on the device (``prep_voxels``, the
``gp_voxel_prep`` kernels, in scipy's operation order and rounding), next to
the decoded field, instead of on the host.

``evaluate_batch`` evaluates a BO batch as one device batch instead of the
reference's four worker threads (:559-611): one VAE decode of all latent
vectors, one pre-processing launch chain, and every solve admitted into the
continuous-batching scheduler together, with GPUVoxelFEM semantics per
request and GPUVoxelFEM's own refinement as scheduler jobs: FP64 residuals
r = f - K u and FP32 Jacobi-PCG (FP64 dot products) corrections K d = r at
``batch_inner_tol`` until each field's residual is below ``batch_tol``, all
on the device. With ``batch_refine = False`` each field gets one FP32 solve
to ``batch_solve_tol`` instead: about twice as fast, compliance within about
1e-5 of the refined value on solid/void fields. Results and history entries
are in parameter order.

With ``use_gpu=False``, or when the GPU solve fails, the evaluation takes
the reference's CalculiX path (:502-557): the C3D8 deck of the solid voxels
(``VoxelHexMesher``, meshed on the device), a ccx run and the .frd parse.
Without a CalculiX executable the result is the reference's no-solver
penalty: stress and compliance FEM_SENTINEL, mass 1.0, ``failure_reason``
"no_ccx" (the deck, which nothing could read, is then not written).
"""
from __future__ import annotations

import json
import logging
from pathlib import Path
from typing import Dict

import numpy as np

from ..pipeline_utils import FEM_SENTINEL, NumpyEncoder

logger = logging.getLogger(__name__)

_NO_CCX = {"stress": FEM_SENTINEL, "compliance": FEM_SENTINEL, "mass": 1.0,
           "failure_reason": "no_ccx"}
_GPU_ERR = {"stress": FEM_SENTINEL, "compliance": FEM_SENTINEL, "mass": 1.0,
            "failure_reason": "gpu_fem_error"}


def _mass(solid_frac: float, shape, bbox: dict | None) -> float:
    D, H, W = shape
    if bbox:
        vol_mm3 = ((bbox["x"][1] - bbox["x"][0]) * (bbox["y"][1] - bbox["y"][0])
                   * (bbox["z"][1] - bbox["z"][0])) * solid_frac
    else:
        vol_mm3 = D * H * W * solid_frac
    return vol_mm3 * 7900 / 1e9


def decode_batch(zs, vae_model):
    """One VAE decode of B latent vectors (voxel_fem.py:449-451 for each):
    (B, D, H, W) float32 fields on the ROCm device."""
    import torch
    zs = [np.asarray(z, dtype=np.float64).reshape(-1) for z in zs]
    device = next(vae_model.parameters()).device
    with torch.no_grad():
        zt = torch.from_numpy(np.stack(zs)).float().to(device)
        dec = vae_model.decode(zt)
    if dec.shape[0] != len(zs):
        raise ValueError(f"decode of {len(zs)} latent vectors returned shape {tuple(dec.shape)}")
    dec = dec.reshape(len(zs), *dec.shape[-3:]).float()
    if not dec.is_cuda:
        dec = dec.to(torch.device("cuda", torch.cuda.current_device()))
    return dec


def decode_to_voxels(zs, vae_model):
    """Decode, 2x down-sampling and the organic density filter
    (voxel_fem.py:449-462) for B latent vectors: (B, d, h, w) float32 device
    tensor."""
    from .voxel_prep import prep_voxels
    return prep_voxels(decode_batch(zs, vae_model))


class VoxelFEMEvaluator:
    """Evaluator with the reference constructor (voxel_fem.py:406-438):
    ``ccx_cmd`` defaults to the CalculiX found on the system, decks of the
    CalculiX path go to ``output_dir``."""

    batch_slots: int = 16           # concurrent solves of one grid size in a batch
    batch_refine: bool = True       # FP64 iterative refinement (GPUVoxelFEM's accuracy)
    batch_tol: float = 1e-8         # refined: FP64 relative residual reached
    batch_inner_tol: float = 1e-3   # refined: FP32 tolerance of every batched solve
    batch_max_outer: int = 60       # refined: correction solves per field at most
    batch_solve_tol: float = 1e-6   # unrefined: tolerance of the one FP32 solve

    def __init__(self, ccx_cmd: str = None, output_dir: str = None, fixed_face: str = "x_min",
                 load_face: str = "x_max", force_n: float = 1000.0, vae_model=None,
                 use_gpu: bool = False):
        from .voxel_hex import find_ccx
        self.ccx_cmd = ccx_cmd or find_ccx()
        self.output_dir = Path(output_dir or "./optimisation_results/voxel_fem")
        self.output_dir.mkdir(parents=True, exist_ok=True)
        self.fixed_face = fixed_face
        self.load_face = load_face
        self.force_n = force_n
        self.vae_model = vae_model
        self.use_gpu = use_gpu
        self.evaluation_history = []
        self._counter = 0
        self._servers = {}

    def close(self):
        for srv in self._servers.values():
            srv.close()
        self._servers = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ccx_path(self, voxels, bbox) -> Dict[str, float]:
        """The reference's CalculiX branch (voxel_fem.py:502-557)."""
        from ..pipeline_utils import is_valid_fem_result
        from .voxel_hex import VoxelHexMesher
        self._counter += 1
        if self.ccx_cmd is None:
            return dict(_NO_CCX)
        inp = str(self.output_dir / f"voxel_fem_{self._counter:04d}.inp")
        try:
            VoxelHexMesher.voxels_to_inp(voxels, bbox=bbox, fixed_face=self.fixed_face,
                                         load_face=self.load_face, force_n=self.force_n,
                                         output_path=inp)
        except ValueError as e:
            logger.warning(f"[VoxelFEM] Mesh failed: {e}")
            return {"stress": FEM_SENTINEL, "compliance": FEM_SENTINEL, "mass": 1.0,
                    "failure_reason": "mesh_error"}
        fem = VoxelHexMesher.run_ccx(inp, ccx_cmd=self.ccx_cmd)
        if not is_valid_fem_result({"stress": fem["stress_max"]}):
            return {"stress": fem["stress_max"], "compliance": fem["compliance"], "mass": 1.0,
                    "failure_reason": fem.get("failure_reason", "fem_failed")}
        v = np.asarray(voxels)
        return {"stress": fem["stress_max"], "compliance": fem["compliance"],
                "mass": _mass(float((v > 0.5).mean()), v.shape, bbox)}

    def evaluate_voxels(self, voxels: np.ndarray, bbox: dict = None) -> Dict[str, float]:
        """FEM metrics of an already decoded and smoothed voxel field."""
        if not self.use_gpu:
            return self._ccx_path(voxels, bbox)
        from .gpu_fem_solver import GPUVoxelFEM
        voxels = np.asarray(voxels)
        try:
            fem = GPUVoxelFEM(voxels=voxels, fixed_face=self.fixed_face, load_face=self.load_face,
                              force_n=self.force_n).solve()
        except Exception as e:  # the reference falls back to ccx here
            logger.warning(f"[VoxelFEM] GPU FEM failed: {e}, falling back to ccx")
            if self.ccx_cmd is not None:
                return self._ccx_path(voxels, bbox)
            self._counter += 1
            return dict(_GPU_ERR)
        result = {"stress": fem["stress_max"], "compliance": fem["compliance"],
                  "mass": _mass(float((voxels > 0.5).mean()), voxels.shape, bbox)}
        self._counter += 1
        self.evaluation_history.append(result)
        return result

    def evaluate(self, z: np.ndarray, vae_model, bbox: dict = None) -> Dict[str, float]:
        """Decode z -> voxels -> FEM -> {stress, compliance, mass}."""
        if not self.use_gpu and self.ccx_cmd is None:  # the penalty needs no field
            self._counter += 1
            return dict(_NO_CCX)
        vox = decode_to_voxels([z], vae_model)[0]
        return self.evaluate_voxels(vox.cpu().numpy(), bbox)

    def _server(self, shape):
        from ..serve import Server, ServeConfig
        srv = self._servers.get(shape)
        if srv is None:
            refine = bool(self.batch_refine)
            cfg = ServeConfig(force_n=float(self.force_n),
                              solve_tol=float(self.batch_tol if refine else self.batch_solve_tol),
                              solve_refine=refine, solve_inner_tol=float(self.batch_inner_tol),
                              solve_max_outer=int(self.batch_max_outer),
                              bcs={"fixed_face": self.fixed_face, "load_face": self.load_face,
                                   "load_dof": 2})
            srv = Server({shape: int(self.batch_slots)}, cfg=cfg, tile=(14, 18), ni=8)
            self._servers[shape] = srv
        return srv

    def evaluate_latents(self, zs, bbox: dict = None) -> list:
        """Result dicts of ``evaluate`` for B latent vectors, evaluated as one
        device batch (decode, pre-processing and solves)."""
        zs = list(zs)
        if not zs:
            return []
        if not self.use_gpu:
            if self.ccx_cmd is None:
                self._counter += len(zs)
                return [dict(_NO_CCX) for _ in zs]
            vox = decode_to_voxels(zs, self.vae_model).cpu().numpy()
            return [self._ccx_path(v, bbox) for v in vox]
        from ..problems import Request
        vox = decode_to_voxels(zs, self.vae_model)
        shape = tuple(int(n) for n in vox.shape[1:])
        solid = (vox > 0.5).reshape(len(zs), -1).float().mean(dim=1).double().cpu().numpy()
        reqs = [Request(b, 0.0, shape, vox[b], "solve", 0.0) for b in range(len(zs))]
        try:
            from ..serve import fem_outputs_batch
            srv = self._server(shape)
            if srv.cfg.solve_refine != bool(self.batch_refine):
                srv.close()
                self._servers.pop(shape)
                srv = self._server(shape)
            # the output formulas run once for the whole batch (two read-backs)
            res = srv.run(reqs, keep_u=True, outputs=False)
            failed = {b for b, r in enumerate(res) if r.status != 0}
            fem = [(r.compliance, r.stress_max) for r in res]
            ok = [b for b in range(len(zs)) if b not in failed]
            outs = fem_outputs_batch([res[b].compliance for b in ok], [res[b].u for b in ok],
                                     vox[ok])
            for b, (c, sm, _) in zip(ok, outs):
                fem[b] = (c, sm)
        except Exception as e:
            logger.warning(f"[VoxelFEM] batched GPU FEM failed: {e}")
            srv = self._servers.pop(shape, None)
            if srv is not None:
                srv.close()
            self._counter += len(zs)
            return [dict(_GPU_ERR) for _ in zs]
        out = []
        for b in range(len(zs)):
            self._counter += 1
            if b in failed:
                logger.warning(f"[VoxelFEM] batched solve {b} failed (status {res[b].status})")
                out.append(dict(_GPU_ERR))
                continue
            result = {"stress": fem[b][1], "compliance": fem[b][0],
                      "mass": _mass(float(solid[b]), shape, bbox)}
            self.evaluation_history.append(result)
            out.append(result)
        return out

    def evaluate_batch(self, param_list: list) -> list:
        """One FEMResult per entry of ``param_list`` (reads ``latent_z``),
        in order (voxel_fem.py:559-611); the entries with a latent vector
        are evaluated as one device batch."""
        if self.vae_model is None:
            raise RuntimeError("VoxelFEMEvaluator.evaluate_batch requires vae_model set at construction")
        from ..schema import FEMResult, OptimizationSample

        results = [None] * len(param_list)
        todo = []
        for i, p in enumerate(param_list):
            z = getattr(p, "latent_z", None)
            if z is None:
                results[i] = FEMResult(stress_max=FEM_SENTINEL, compliance=FEM_SENTINEL, mass=1.0,
                                       success=False)
            else:
                todo.append((i, z))
        dicts = self.evaluate_latents([z for _, z in todo])
        for (i, z), r in zip(todo, dicts):
            res = FEMResult(stress_max=float(r.get("stress", FEM_SENTINEL)),
                            compliance=float(r.get("compliance", FEM_SENTINEL)),
                            mass=float(r.get("mass", 1.0)),
                            success=r.get("failure_reason") is None)
            results[i] = res
            self.evaluation_history.append(OptimizationSample(
                parameters=param_list[i], result=res,
                latent_z=z.tolist() if hasattr(z, "tolist") else list(z)))
        return results

    def save_history(self, path: str):
        with open(path, "w") as f:
            json.dump(self.evaluation_history, f, indent=2, cls=NumpyEncoder)
