"""SIMP solvers, the data generator and the STL export with the interfaces of
genpipeline/topology (simp_solver.py, simp_solver_gpu.py,
simp_solver_dense_gpu.py, topo_data_gen.py, mesh_export.py), running on the
HIP library."""
from .mesh_export import density_to_stl
from .simp_solver import SIMPSolver
from .simp_solver_dense_gpu import DenseGPUSolver
from .simp_solver_gpu import SIMPSolverGPU, StiffnessOperator
from .topo_data_gen import TopoDataGenerator

__all__ = ["SIMPSolver", "SIMPSolverGPU", "DenseGPUSolver", "StiffnessOperator",
           "TopoDataGenerator", "density_to_stl"]
