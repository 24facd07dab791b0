"""Result and parameter records exchanged with the evaluator
(genpipeline/schema.py:5-35): the fields the FEM path reads and writes."""
from __future__ import annotations

from typing import Any, Dict, List, Optional

from pydantic import BaseModel, ConfigDict, Field


class DesignParameters(BaseModel):
    """schema.py:5-14."""
    model_config = ConfigDict(arbitrary_types_allowed=True)

    h_mm: float = Field(..., description="Primary height/dimension in mm")
    r_mm: float = Field(..., description="Secondary radius/thickness in mm")
    geometry_type: str = Field("cantilever", description="Geometry type")
    material_name: str = Field("Plastic_ABS", description="Material identifier")
    material_cfg: Optional[Dict[str, Any]] = Field(None, description="Material properties")
    latent_z: Optional[List[float]] = Field(None, description="Latent vector")


class FEMResult(BaseModel):
    """schema.py:16-25."""
    model_config = ConfigDict(arbitrary_types_allowed=True)

    stress_max: float = Field(..., description="Maximum stress (MPa)")
    stress_mean: Optional[float] = Field(None, description="Mean stress (MPa)")
    compliance: float = Field(..., description="Structural compliance")
    mass: float = Field(..., description="Mass (kg)")
    bbox: Optional[Dict[str, float]] = Field(None, description="Bounding box")
    success: bool = Field(True, description="Whether the evaluation succeeded")


class OptimizationSample(BaseModel):
    """schema.py:27-34."""
    model_config = ConfigDict(arbitrary_types_allowed=True)

    parameters: Any
    result: FEMResult
    latent_z: Optional[List[float]] = Field(None, description="Latent vector")
    timestamp: Optional[float] = Field(None, description="Unix timestamp")
