"""Voxel FEM evaluation with the interfaces of genpipeline/fem
(gpu_fem_solver.py, voxel_fem.py GPU branch), running on the HIP library."""
from .gpu_fem_solver import GPUConjugateGradientFEM, GPUVoxelFEM
from .voxel_fem import VoxelFEMEvaluator

__all__ = ["GPUVoxelFEM", "GPUConjugateGradientFEM", "VoxelFEMEvaluator"]
