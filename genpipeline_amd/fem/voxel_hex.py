"""CalculiX cross-check path: voxel field -> C3D8 hex mesh (.inp), ccx run,
.frd parse. Drop-in for ``VoxelHexMesher`` and ``find_ccx``
(genpipeline/fem/voxel_fem.py:38-399) and the solid-voxel scan
``get_solid_voxels_simd`` (genpipeline/cuda_kernels/__init__.py:371-388).

The mesh is built on the device (``gp_hex_mesh_*``): one C3D8 element per
voxel > 0.5 in C order, grid corners numbered from 1 in the order the
elements first touch them (corners in C3D8 order), as the reference's
node-table loop numbers them (voxel_fem.py:120-146). The host writes the
.inp text in the reference's layout: node and element cards, material and
section, NFIX / NLOAD node sets in node-number order (16 a line), the
clamp, the load -force_n / n_load_nodes in -z on every load-face node, and
the U / S output requests (:181-225). ``run_ccx`` runs CalculiX when one is
installed and parses its .frd with the reference's fixed-width reader
(:333-399): max nodal von Mises stress, max displacement magnitude, and as
"compliance" the sum of displacement magnitudes, with the reference's
sentinels on every failure.

No ``ccx`` binary exists in this image, so the cross-check itself runs
only where CalculiX is installed; the mesh and file formats are tested
here against the reference's own fixtures.
"""
from __future__ import annotations

import ctypes as C
import logging
import multiprocessing
import os
import shutil
import subprocess
import threading
from pathlib import Path
from typing import Dict, Optional

import numpy as np
import torch

from .. import _lib as L
from ..pipeline_utils import FEM_SENTINEL

logger = logging.getLogger(__name__)

# concurrent ccx processes, as the reference limits them (voxel_fem.py:35)
_CCX_SEMAPHORE = threading.Semaphore(max(1, multiprocessing.cpu_count() // 2))

FACES = ("x_min", "x_max", "y_min", "y_max", "z_min", "z_max")


def find_ccx() -> Optional[str]:
    """The CalculiX executable: $CCX_PATH, else ``ccx`` on PATH, else None
    (voxel_fem.py:38-68)."""
    for p in [os.environ.get("CCX_PATH"), shutil.which("ccx")]:
        if p and (shutil.which(p) or Path(p).exists()):
            return p
    logger.warning("[VoxelFEM] CalculiX (ccx) not found on this system.")
    return None


def _device_field(voxels) -> torch.Tensor:
    if not torch.cuda.is_available():
        raise RuntimeError("the voxel mesh is built on a ROCm GPU (torch.cuda.is_available() is False)")
    if isinstance(voxels, torch.Tensor):
        t = voxels.detach()
        if not t.is_cuda:
            t = t.cuda()
    else:
        t = torch.from_numpy(np.ascontiguousarray(voxels, dtype=np.float32)).cuda()
    t = t.to(torch.float32).contiguous()
    if t.dim() != 3:
        raise ValueError("voxels must be a 3-D array")
    return t


def hex_mesh(voxels):
    """C3D8 mesh of the voxels > 0.5 (compared in float32, as the reference's
    SIMD scanner compares). Returns (elements (n_elem, 8) int32 node numbers
    from 1, node_ijk (n_node, 3) int32 grid corners in node-number order,
    elem_voxel (n_elem,) int32 C index of each element's voxel), host arrays."""
    t = _device_field(voxels)
    D, H, W = (int(n) for n in t.shape)
    n_ints = C.c_int64()
    L.call("gp_hex_mesh_workspace", D, H, W, C.byref(n_ints))
    dev = t.device
    ws = torch.empty(int(n_ints.value), dtype=torch.int32, device=dev)
    ne, nn = C.c_int64(), C.c_int64()
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    with torch.cuda.device(dev):
        L.call("gp_hex_mesh_count", C.c_void_p(t.data_ptr()), D, H, W, C.c_void_p(ws.data_ptr()),
               C.byref(ne), C.byref(nn), stream)
        elems = torch.empty((max(ne.value, 1), 8), dtype=torch.int32, device=dev)
        ijk = torch.empty((max(nn.value, 1), 3), dtype=torch.int32, device=dev)
        vox = torch.empty(max(ne.value, 1), dtype=torch.int32, device=dev)
        L.call("gp_hex_mesh_emit", C.c_void_p(t.data_ptr()), D, H, W, C.c_void_p(ws.data_ptr()),
               C.c_void_p(elems.data_ptr()), C.c_void_p(ijk.data_ptr()),
               C.c_void_p(vox.data_ptr()), stream)
    n_e, n_n = int(ne.value), int(nn.value)
    return (elems[:n_e].cpu().numpy(), ijk[:n_n].cpu().numpy(), vox[:n_e].cpu().numpy())


def get_solid_voxels(voxels) -> np.ndarray:
    """(n, 3) int64 indices of the voxels > 0.5 in C order: the reference's
    get_solid_voxels_simd (cuda_kernels/__init__.py:371-388), on the device."""
    t = _device_field(voxels)
    _, H, W = (int(n) for n in t.shape)
    _, _, vox = hex_mesh(t)
    v = vox.astype(np.int64)
    return np.stack([v // (H * W), (v // W) % H, v % W], axis=1)


def _face_mask(face, ijk, D, H, W):
    cx, cy, cz = ijk[:, 0], ijk[:, 1], ijk[:, 2]
    return {"x_min": cx == 0, "x_max": cx == D, "y_min": cy == 0, "y_max": cy == H,
            "z_min": cz == 0, "z_max": cz == W}.get(face, np.zeros(len(ijk), bool))


def _id_lines(ids):
    return [",".join(str(n) for n in ids[i:i + 16]) for i in range(0, len(ids), 16)]


class VoxelHexMesher:
    """Voxel grid -> CalculiX C3D8 deck, ccx run, .frd parse
    (voxel_fem.py:74-399)."""

    @staticmethod
    def voxels_to_inp(voxels: np.ndarray, bbox: Optional[Dict] = None, fixed_face: str = "x_min",
                      load_face: str = "x_max", force_n: float = 1000.0,
                      E_mpa: float = 210000.0, poisson: float = 0.30,
                      output_path: str = "voxel_fem.inp") -> str:
        """Write the .inp deck of the solid voxels (voxel_fem.py:76-226);
        returns its path. ValueError for an all-void field or a face with
        no mesh node, as the reference raises."""
        arr = voxels if isinstance(voxels, torch.Tensor) else np.asarray(voxels)
        D, H, W = (int(n) for n in arr.shape)
        if not bool((arr > 0.5).any()):
            raise ValueError("All voxels are void — cannot mesh empty geometry.")
        if bbox is None:
            dx = dy = dz = 1.0
            x0 = y0 = z0 = 0.0
        else:
            x0, x1 = bbox["x"]
            y0, y1 = bbox["y"]
            z0, z1 = bbox["z"]
            dx, dy, dz = (x1 - x0) / D, (y1 - y0) / H, (z1 - z0) / W
        elems, ijk, _ = hex_mesh(arr)
        if len(elems) == 0:
            raise ValueError("No solid voxels found to mesh.")
        ids = np.arange(1, len(ijk) + 1)
        fixed = ids[_face_mask(fixed_face, ijk, D, H, W)].tolist()
        load = ids[_face_mask(load_face, ijk, D, H, W)].tolist()
        if not fixed:
            raise ValueError(f"No nodes on fixed face '{fixed_face}'.")
        if not load:
            raise ValueError(f"No nodes on load face '{load_face}'.")
        force_per_node = force_n / len(load)
        lines = ["** CalculiX C3D8 hex mesh — generated by voxel_fem.py", "*NODE, NSET=NALL"]
        for nid, (cx, cy, cz) in enumerate(ijk.tolist(), 1):
            lines.append(f"{nid},{x0 + cx * dx:.6f},{y0 + cy * dy:.6f},{z0 + cz * dz:.6f}")
        lines.append("*ELEMENT, TYPE=C3D8, ELSET=EALL")
        for eid, en in enumerate(elems.tolist(), 1):
            lines.append(f"{eid}," + ",".join(str(n) for n in en))
        lines += ["*MATERIAL, NAME=STEEL", "*ELASTIC", f"{E_mpa},{poisson}",
                  "*SOLID SECTION, ELSET=EALL, MATERIAL=STEEL", "**", "*NSET, NSET=NFIX"]
        lines += _id_lines(fixed)
        lines.append("*NSET, NSET=NLOAD")
        lines += _id_lines(load)
        lines += ["**", "*STEP", "*STATIC", "*BOUNDARY", "NFIX,1,3,0.0", "*CLOAD"]
        lines += [f"{nid},3,{-force_per_node:.6f}" for nid in load]
        lines += ["*NODE FILE", "U", "*EL FILE", "S", "*END STEP"]
        output_path = str(output_path)
        with open(output_path, "w") as f:
            f.write("\n".join(lines) + "\n")
        logger.info(f"[VoxelFEM] Wrote .inp: {output_path} ({len(ijk)} nodes, {len(elems)} "
                    f"elements, {len(fixed)} fixed nodes, {len(load)} load nodes)")
        return output_path

    @staticmethod
    def run_ccx(inp_path: str, ccx_cmd: str = None, timeout: int = 300) -> Dict[str, float]:
        """Run CalculiX on a deck (voxel_fem.py:228-302): the parsed .frd, or
        sentinels with failure_reason ccx_error / timeout / no_frd."""
        if ccx_cmd is None:
            ccx_cmd = find_ccx()
        if ccx_cmd is None:
            raise RuntimeError("CalculiX (ccx) not found. Set ccx_cmd explicitly.")
        work_dir = Path(inp_path).parent
        stem = Path(inp_path).stem
        arg = str(work_dir / stem) if ccx_cmd.endswith(".exe") else stem
        sentinel = {"stress_max": FEM_SENTINEL, "displacement_max": FEM_SENTINEL,
                    "compliance": FEM_SENTINEL}
        with _CCX_SEMAPHORE:
            try:
                proc = subprocess.run([ccx_cmd, arg], capture_output=True, text=True,
                                      timeout=timeout, cwd=str(work_dir))
                if proc.returncode != 0:
                    logger.warning(f"[VoxelFEM] ccx returned non-zero exit code {proc.returncode}")
                    return {**sentinel, "failure_reason": "ccx_error"}
            except subprocess.TimeoutExpired:
                logger.error(f"[VoxelFEM] ccx timed out after {timeout}s")
                return {**sentinel, "failure_reason": "timeout"}
        frd = work_dir / f"{stem}.frd"
        if not frd.exists():
            return {**sentinel, "failure_reason": "no_frd"}
        out = VoxelHexMesher._parse_frd(str(frd))
        out["failure_reason"] = None
        return out

    @staticmethod
    def _parse_frd(frd_path: str) -> Dict[str, float]:
        """Max von Mises stress, max displacement magnitude and the sum of
        displacement magnitudes from a .frd (voxel_fem.py:333-399). Data
        records ' -1' carry a 10-character node number after the 3-character
        record code, then 12-character values; DISP and STRESS blocks open
        with '-4' and close with '-3' or '-2'. Unreadable records are
        skipped; an unreadable file gives the sentinels."""
        disps, stresses = [], []
        in_disp = in_stress = False

        def val(line, slot):
            a = 13 + 12 * slot
            return float(line[a:a + 12])

        try:
            with open(frd_path) as f:
                for line in f:
                    st = line.strip()
                    if "-4  DISP" in line or "-4 DISP" in line:
                        in_disp, in_stress = True, False
                        continue
                    if "-4  STRESS" in line or "-4 STRESS" in line:
                        in_stress, in_disp = True, False
                        continue
                    if st.startswith("-3") or st.startswith("-2"):
                        in_disp = in_stress = False
                        continue
                    if st.startswith("-5") or st.startswith("-4"):
                        continue
                    if not (line.startswith(" -1") and len(line) > 25):
                        continue
                    try:
                        if in_disp:
                            ux, uy, uz = val(line, 0), val(line, 1), val(line, 2)
                            disps.append((ux ** 2 + uy ** 2 + uz ** 2) ** 0.5)
                        elif in_stress and len(line) >= 85:
                            s11, s22, s33 = val(line, 0), val(line, 1), val(line, 2)
                            s12, s23, s13 = val(line, 3), val(line, 4), val(line, 5)
                            stresses.append((0.5 * ((s11 - s22) ** 2 + (s22 - s33) ** 2
                                                    + (s33 - s11) ** 2
                                                    + 6 * (s12 ** 2 + s23 ** 2 + s13 ** 2))) ** 0.5)
                    except (ValueError, IndexError):
                        pass
        except OSError as e:
            logger.error(f"[VoxelFEM] Could not read .frd: {e}")
            return {"stress_max": FEM_SENTINEL, "displacement_max": FEM_SENTINEL,
                    "compliance": FEM_SENTINEL, "failure_reason": "no_frd"}
        return {"stress_max": float(max(stresses)) if stresses else 0.0,
                "displacement_max": float(max(disps)) if disps else 0.0,
                "compliance": float(sum(disps)) if disps else 0.0}
