"""Voxel FEM evaluation with the interfaces of genpipeline/fem
(gpu_fem_solver.py, voxel_fem.py: GPU branch and CalculiX path), running on
the HIP library."""
from .gpu_fem_solver import GPUConjugateGradientFEM, GPUVoxelFEM
from .voxel_fem import VoxelFEMEvaluator
from .voxel_hex import VoxelHexMesher, find_ccx, get_solid_voxels

__all__ = ["GPUVoxelFEM", "GPUConjugateGradientFEM", "VoxelFEMEvaluator", "VoxelHexMesher",
           "find_ccx", "get_solid_voxels"]
