"""FEM result sentinels, the JSON encoder and the post-decode voxel filter
that the evaluator path uses (genpipeline/pipeline_utils.py:20-76).

Same names, defaults and return conventions as the reference module, so
callers that test ``is_valid_fem_result`` or compare against ``FEM_SENTINEL``
keep working.
"""
from __future__ import annotations

import json

import numpy as np

# pipeline_utils.py:20-21
FEM_SENTINEL: float = 1e6
FEM_VALID_THRESHOLD: float = 1e5


def is_valid_fem_result(result) -> bool:
    """True when ``result`` holds a plausible (non-sentinel) stress:
    0 < stress < FEM_VALID_THRESHOLD (pipeline_utils.py:24-33). Accepts a
    FEMResult, or a dict with "stress" or "stress_max"."""
    from .schema import FEMResult
    if isinstance(result, FEMResult):
        stress = result.stress_max
    elif isinstance(result, dict):
        stress = result.get("stress", result.get("stress_max", FEM_SENTINEL))
    else:
        stress = FEM_SENTINEL
    return 0.0 < float(stress) < FEM_VALID_THRESHOLD


class NumpyEncoder(json.JSONEncoder):
    """JSON encoder for numpy scalars/arrays and pydantic models."""

    def default(self, obj):
        if isinstance(obj, np.ndarray):
            return obj.tolist()
        if isinstance(obj, np.integer):
            return int(obj)
        if isinstance(obj, np.floating):
            return float(obj)
        dump = getattr(obj, "model_dump", None)
        if callable(dump):
            return dump()
        return super().default(obj)


def smooth_voxels(voxels: np.ndarray, sigma: float = 0.5, sharpness: float = 15.0) -> np.ndarray:
    """Gaussian blur (``sigma`` voxels) then logistic sharpening
    1 / (1 + exp(-sharpness (v - 0.5))) (pipeline_utils.py:55-76)."""
    from scipy.ndimage import gaussian_filter

    v = gaussian_filter(np.asarray(voxels), sigma=sigma)
    return 1.0 / (1.0 + np.exp(-sharpness * (v - 0.5)))
