"""TopoDataGenerator: SIMP runs as a dataset generator, batched on the device.

Drop-in for genpipeline/topology/topo_data_gen.py:37-340 (class tables,
constructor, ``generate_single``, ``generate``, ``generate_scaled``). A
sample is drawn as the reference draws it, from NumPy's global generator in
the same order (geometry by weight, grid size by class, volume fraction and
force; :168-186, :235-244); solved by SIMP; and written as the reference
writes it: ``{stem}_mesh.stl`` (density_to_stl at 0.5 with voxel size
(100/nx, 20/ny, 20/nz) mm) and ``{stem}_fem_results.json`` with
stress_max = 0.1 compliance, compliance, mass = mean density x volume x
1.05e-6 and the parameters, dumped with indent 2 (:246-291).

Solver semantics follow the reference's choice (:152-166):

* ``use_gpu=True`` (default): DenseGPUSolver semantics -- supports on the
  x_min, y_min and z_min faces, the load spread over the x_max face, the
  geometry's boundary conditions unused; ``compliance`` is the solve before
  the last OC update;
* ``use_gpu=False``: SIMPSolver semantics with the geometry's boundary
  conditions; ``compliance`` is the solve after the last update.

``generate`` and ``generate_scaled`` draw all samples of a batch first and
run their SIMP problems together in the device scheduler (independent
progress, one tick launch for all) instead of one solver per worker thread;
``generate_single`` runs one sample through the solver facade, as the
reference does. In the batch the PCG is FP32 with FP64 dot products to the
solver's tolerance (DenseGPUSolver's 1e-5; 1e-6 where the SIMPSolver facade
refines to 1e-10), so batched densities and compliances agree with the
facades' to that tolerance, not bit for bit. Samples are drawn in index
order, so a seeded generator reproduces its parameter sequence (the
reference's thread pool draws in completion order).
"""
from __future__ import annotations

import json
import logging
import uuid
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .mesh_export import density_to_stl

log = logging.getLogger(__name__)


@dataclass
class Sample:
    geom: str
    bcs: dict
    nx: int
    ny: int
    nz: int
    volfrac: float
    force_mag: float


class TopoDataGenerator:
    """Generates topology optimisation samples by SIMP (topo_data_gen.py:37)."""

    # boundary conditions per geometry (topo_data_gen.py:40-95)
    GEOM_BCS = {
        "cantilever": {"fixed_face": "x_min", "load_face": "x_max", "load_dof": 2},
        "tapered": {"fixed_face": "x_min", "load_face": "x_max", "load_dof": 2},
        "ribbed": {"fixed_face": "x_min", "load_face": "x_max", "load_dof": 2},
        "lbracket": {"fixed_face": "z_min", "load_face": "x_max", "load_dof": 2},
        "bridge": {"fixed_face": ["x_min", "x_max"], "load_face": "center", "load_dof": 2},
        "arch": {"fixed_face": ["x_min", "x_max"], "load_face": "top", "load_dof": 2},
        "3dload": {"fixed_face": "x_min", "load_face": "x_max", "load_dof": [1, 2],
                   "multi_load": True},
        "offset": {"fixed_face": "x_min", "load_face": "x_max_offset", "load_dof": 2},
        "center": {"fixed_face": "x_min", "load_face": "center_z", "load_dof": 2},
        "portal": {"fixed_face": ["y_min", "y_max"], "load_face": "top_center", "load_dof": 2},
        "spoke": {"fixed_face": "center", "load_face": "outer_ring", "load_dof": 2},
        "tower": {"fixed_face": "z_min", "load_face": "z_max", "load_dof": [1, 2]},
        "sandwich": {"fixed_face": ["z_min", "z_max"], "load_face": "center", "load_dof": 2},
        "simply": {"fixed_face": ["x_min", "x_max"], "load_face": "top", "load_dof": 2},
        "branch": {"fixed_face": "x_min", "load_face": "x_max", "load_dof": 2,
                   "multi_load": True, "load_positions": ["top", "mid", "bottom"]},
        "y_tree": {"fixed_face": "z_min", "load_face": "x_max", "load_dof": 2,
                   "multi_load": True, "load_positions": ["top", "mid_y", "mid_z"]},
        "network": {"fixed_face": ["x_min", "y_min"], "load_face": "x_max", "load_dof": 2,
                    "multi_load": True, "load_pattern": "grid"},
    }

    # sampling weights in organic mode; unlisted geometries weigh 1 (the
    # class body's second, effective table, topo_data_gen.py:112-125)
    GEOM_WEIGHTS = {
        "bridge": 2, "arch": 2, "3dload": 2, "offset": 1, "center": 1, "cantilever": 1,
        "tapered": 1, "ribbed": 1, "lbracket": 1, "branch": 1, "y_tree": 1, "network": 1,
    }

    def __init__(self, output_dir: str = "./fem_data", n_workers: int = 2, use_gpu: bool = True,
                 n_iters: int = 30, grid_size: str = "medium", penal: float = 3.5,
                 organic_mode: bool = True):
        self.output_dir = Path(output_dir)
        self.output_dir.mkdir(parents=True, exist_ok=True)
        self.n_workers = n_workers
        self.use_gpu = use_gpu
        self.n_iters = n_iters
        self.grid_size = grid_size
        self.penal = penal
        self.organic_mode = organic_mode
        self._SIMP_GPU = None
        self._SIMP_CPU = None

    # -------------------------------------------------------------- draws
    def _get_grid_dims(self, geom: str):
        """Grid by size class and geometry (topo_data_gen.py:180-221)."""
        ri = np.random.randint
        if self.grid_size == "small":
            return 16, 8, 8
        if self.grid_size == "medium":
            if geom in ("lbracket", "bridge", "arch", "center"):
                return ri(16, 24), ri(6, 10), ri(12, 18)
            if geom in ("3dload", "offset"):
                return ri(22, 32), ri(8, 12), ri(8, 12)
            if geom in ("branch", "y_tree", "network"):
                return ri(22, 32), ri(10, 16), ri(10, 16)
            return ri(20, 32), ri(6, 10), ri(6, 10)
        if geom in ("lbracket", "bridge", "arch", "center"):
            return ri(24, 36), ri(8, 12), ri(24, 36)
        if geom in ("3dload", "offset"):
            return ri(28, 40), ri(8, 14), ri(8, 14)
        if geom in ("branch", "y_tree", "network"):
            return ri(24, 36), ri(12, 18), ri(12, 18)
        return ri(24, 40), ri(6, 12), ri(6, 12)

    def _create_solver(self, nx, ny, nz, bcs):
        """DenseGPUSolver or SIMPSolver facade (topo_data_gen.py:223-239)."""
        if self.use_gpu:
            if self._SIMP_GPU is None:
                from .simp_solver_dense_gpu import DenseGPUSolver
                self._SIMP_GPU = DenseGPUSolver
            return self._SIMP_GPU(nx=nx, ny=ny, nz=nz, penal=self.penal, rmin=1.5, device="cuda")
        if self._SIMP_CPU is None:
            from .simp_solver import SIMPSolver
            self._SIMP_CPU = SIMPSolver
        return self._SIMP_CPU(nx=nx, ny=ny, nz=nz, penal=self.penal, rmin=1.5,
                              boundary_conditions=bcs)

    def _select_geometry(self):
        """Geometry drawn by weight in organic mode, uniformly otherwise
        (topo_data_gen.py:241-248)."""
        geoms = list(self.GEOM_BCS.keys())
        if self.organic_mode:
            weights = [self.GEOM_WEIGHTS.get(g, 1) for g in geoms]
            return np.random.choice(geoms, p=np.array(weights) / sum(weights))
        return np.random.choice(geoms)

    def draw_sample(self) -> Sample:
        """One sample's parameters, drawn in generate_single's order
        (topo_data_gen.py:250-264)."""
        geom = str(self._select_geometry())
        nx, ny, nz = (int(v) for v in self._get_grid_dims(geom))
        if self.organic_mode:
            volfrac = np.random.uniform(0.12, 0.35)
            force_mag = np.random.uniform(200, 5000)
        else:
            volfrac = np.random.uniform(0.2, 0.5)
            force_mag = np.random.uniform(500, 2000)
        return Sample(geom, self.GEOM_BCS[geom], nx, ny, nz, float(volfrac), float(force_mag))

    # ------------------------------------------------------------ outputs
    def _write(self, s: Sample, density: np.ndarray, compliance: float, i: int, n_samples: int):
        """STL and JSON of one sample (topo_data_gen.py:268-300)."""
        sample_id = str(uuid.uuid4())[:8]
        stem = (f"{s.geom[:4]}_v{int(s.volfrac * 100)}_f{int(s.force_mag)}_"
                f"{s.nx}x{s.ny}x{s.nz}_{sample_id}")
        stl_path = self.output_dir / f"{stem}_mesh.stl"
        json_path = self.output_dir / f"{stem}_fem_results.json"
        voxel_size_mm = (100.0 / s.nx, 20.0 / s.ny, 20.0 / s.nz)
        if density_to_stl(density, str(stl_path), threshold=0.5, voxel_size_mm=voxel_size_mm) is None:
            raise ValueError(f"Empty design for {s.geom}")
        total_vol = 100.0 * 20.0 * 20.0
        if s.geom in ("lbracket", "bridge", "arch"):
            total_vol = 100.0 * 20.0 * 100.0
        mass = density.mean() * total_vol * 1.05e-6
        results = {
            "stress_max": float(compliance * 0.1),
            "compliance": float(compliance),
            "mass": float(mass),
            "parameters": {"geometry": s.geom, "volfrac": float(s.volfrac),
                           "force_n": float(s.force_mag), "nx": int(s.nx), "ny": int(s.ny),
                           "nz": int(s.nz), "penal": self.penal},
        }
        with open(json_path, "w") as f:
            json.dump(results, f, indent=2)
        log.info(f"[{i + 1}/{n_samples}] Generated {s.geom}: {stl_path.name}")
        return stem

    # ------------------------------------------------------------ solving
    def generate_single(self, i, n_samples):
        """Draw, solve (one solver facade) and write one sample."""
        s = self.draw_sample()
        solver = self._create_solver(s.nx, s.ny, s.nz, s.bcs)
        density = solver.run(volfrac=s.volfrac, n_iters=self.n_iters, force_mag=s.force_mag)
        return self._write(s, density, solver.last_compliance, i, n_samples)

    def solve_batch(self, samples: list) -> list:
        """(density, compliance) of every sample, all SIMP runs in one device
        scheduler."""
        import torch
        from ..bcs import simp_bcs
        from ..problem import Problem, ProblemSpec
        from ..sched import run_batch
        from .simp_solver_dense_gpu import dense_bcs

        if not samples:
            return []
        dense = self.use_gpu
        problems = []
        for s in samples:
            if dense:
                f, fixed = dense_bcs(s.nx, s.ny, s.nz, s.force_mag)
            else:
                f, fixed = simp_bcs(s.bcs, s.nx, s.ny, s.nz, s.force_mag)
            spec = ProblemSpec(s.nx, s.ny, s.nz, penal=float(self.penal), rmin=1.5,
                               volfrac=s.volfrac, law="clamp", oc="cpu")
            problems.append(Problem(spec, fixed, f))
        try:
            n = int(self.n_iters)
            if n <= 0:
                return [(np.full((s.nx, s.ny, s.nz), s.volfrac), 0.0) for s in samples]
            if dense:
                # n-1 iterations, then the solve before the last update (the
                # job's final solve) and the last update itself
                res, _ = run_batch(problems, n - 1, [s.volfrac for s in samples], tile=(14, 18),
                                   graph_ticks=32, tol=1e-5, max_iter=20000, warm_start=False)
                out = []
                for s, P, r in zip(samples, problems, res):
                    dc = P.filter(P.sens(r.xPhys, r.u.to(torch.float32)), mode=1)
                    _, xPhys, _, _ = P.oc(r.x, dc, volfrac=s.volfrac)
                    out.append((xPhys.float().cpu().numpy().reshape(s.nx, s.ny, s.nz),
                                float(r.compliance)))
                return out
            res, _ = run_batch(problems, n, [s.volfrac for s in samples], tile=(14, 18),
                               graph_ticks=32, tol=1e-6, max_iter=20000, warm_start=True)
            return [(r.xPhys.cpu().numpy().reshape(s.nx, s.ny, s.nz), float(r.compliance))
                    for s, r in zip(samples, res)]
        finally:
            torch.cuda.synchronize()
            for P in problems:
                P.close()

    def _generate_range(self, first: int, count: int, n_samples: int) -> list:
        samples = [self.draw_sample() for _ in range(count)]
        stems = []
        for k, (s, (density, compliance)) in enumerate(zip(samples, self.solve_batch(samples))):
            try:
                stems.append(self._write(s, density, compliance, first + k, n_samples))
            except Exception as e:  # the reference logs a failed sample and goes on
                log.error(f"Sample generation failed: {e}")
        return stems

    def generate(self, n_samples: int = 100):
        """n_samples samples as one device batch (topo_data_gen.py:302-317);
        returns the stems written."""
        log.info(f"Generating {n_samples} samples as one device batch...")
        return self._generate_range(0, n_samples, n_samples)

    def generate_scaled(self, n_samples: int = 1000, batch_size: int = 100):
        """n_samples in device batches of batch_size (topo_data_gen.py:319-340)."""
        log.info(f"Generating {n_samples} samples in batches of {batch_size}...")
        stems = []
        total_batches = (n_samples + batch_size - 1) // batch_size
        generated = 0
        for batch_idx in range(total_batches):
            count = min(batch_size, n_samples - generated)
            log.info(f"Batch {batch_idx + 1}/{total_batches}: Generating {count} samples...")
            stems += self._generate_range(generated, count, n_samples)
            generated += count
            log.info(f"Progress: {generated}/{n_samples} samples generated")
        return stems
