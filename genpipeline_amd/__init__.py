"""MI355X-native SIMP topology-optimisation inner loop (hex8 K.u, Jacobi-PCG,
sensitivity, density filter, OC) behind genpipeline's solver signatures."""
from .problem import Problem, ProblemSpec, footprint_bytes  # noqa: F401

__all__ = ["Problem", "ProblemSpec", "footprint_bytes"]
