"""SIMP solvers with the interfaces of genpipeline/topology (simp_solver.py,
simp_solver_gpu.py, simp_solver_dense_gpu.py), running on the HIP library."""
from .simp_solver import SIMPSolver
from .simp_solver_dense_gpu import DenseGPUSolver
from .simp_solver_gpu import SIMPSolverGPU, StiffnessOperator

__all__ = ["SIMPSolver", "SIMPSolverGPU", "DenseGPUSolver", "StiffnessOperator"]
