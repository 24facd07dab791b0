"""density_to_stl: a density field's 0.5 isosurface as a binary STL file.

Drop-in for genpipeline/topology/mesh_export.py:7-67 (same name, signature,
defaults and return / error behaviour). The reference runs
skimage.measure.marching_cubes on the host and writes the faces with
numpy-stl; here the marching cubes run on the device (``gp_mc_count`` /
``gp_mc_emit``) and the records are written directly:

* an empty design (max < 0.01) returns None;
* a threshold outside [min, max] is moved 10 % into the range, and the
  level is at least 0.1 (mesh_export.py:38-49);
* no surface at that level raises ValueError (mesh_export.py:51-61);
* vertices are in index units times ``voxel_size_mm``.

The triangulation is this library's own (a generated marching-cubes table;
vertices at the linear level crossing of each grid edge, as skimage places
them), so the triangle lists are not skimage's. The file is standard binary
STL: an 80-byte header, the triangle count, then per triangle the unit
normal, three vertices (float32) and a zero attribute word.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib as L

_HEADER = b"genpipeline_amd density_to_stl".ljust(80, b" ")
_RECORD = np.dtype([("nv", "<f4", (12,)), ("attr", "<u2")])


def _level(density: np.ndarray, threshold: float):
    if float(density.max()) < 0.01:
        return None
    lo, hi = float(density.min()), float(density.max())
    if threshold < lo:
        threshold = lo + 0.1 * (hi - lo)
    if threshold > hi:
        threshold = hi - 0.1 * (hi - lo)
    return max(threshold, 0.1)


def marching_cubes(density, level: float, scale=(1.0, 1.0, 1.0), device=None) -> torch.Tensor:
    """(T, 12) float32 device tensor of STL records (unit normal, then three
    vertices) for the surface density = level of a 3-D field."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    d = torch.as_tensor(density, dtype=torch.float64, device=dev).contiguous()
    if d.dim() != 3:
        raise ValueError(f"expected a 3-D density field, got shape {tuple(d.shape)}")
    nx, ny, nz = (int(s) for s in d.shape)
    n_ws = C.c_int64()
    L.call("gp_mc_workspace", nx, ny, nz, C.byref(n_ws))
    ws = torch.empty(n_ws.value, dtype=torch.int32, device=dev)
    n_tri = C.c_int64()
    with torch.cuda.device(dev):
        st = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
        L.call("gp_mc_count", C.c_void_p(d.data_ptr()), nx, ny, nz, C.c_double(level),
               C.c_void_p(ws.data_ptr()), C.byref(n_tri), st)
        out = torch.empty((n_tri.value, 12), dtype=torch.float32, device=dev)
        if n_tri.value:
            sx, sy, sz = (float(s) for s in scale)
            L.call("gp_mc_emit", C.c_void_p(d.data_ptr()), nx, ny, nz, C.c_double(level),
                   C.c_double(sx), C.c_double(sy), C.c_double(sz), C.c_void_p(ws.data_ptr()),
                   C.c_void_p(out.data_ptr()), st)
    return out


def write_stl(path: str, records: np.ndarray) -> None:
    """Binary STL of (T, 12) float32 records."""
    rec = np.zeros(len(records), dtype=_RECORD)
    rec["nv"] = np.asarray(records, np.float32)
    with open(path, "wb") as f:
        f.write(_HEADER)
        f.write(np.uint32(len(rec)).tobytes())
        f.write(rec.tobytes())


def read_stl(path: str) -> np.ndarray:
    """(T, 12) float32 records of a binary STL file."""
    with open(path, "rb") as f:
        f.read(80)
        n = int(np.frombuffer(f.read(4), "<u4")[0])
        rec = np.frombuffer(f.read(n * _RECORD.itemsize), dtype=_RECORD)
    return np.array(rec["nv"])


def density_to_stl(density: np.ndarray, output_path: str, threshold: float = 0.5,
                   voxel_size_mm: tuple = (100 / 32, 20 / 8, 20 / 8)) -> str:
    """Write the density field's isosurface to ``output_path`` as STL
    (mesh_export.py:7-67); returns the path, or None for an empty design."""
    d = np.asarray(density)
    level = _level(d, threshold)
    if level is None:
        return None
    rec = marching_cubes(d, level, voxel_size_mm)
    if rec.shape[0] == 0:
        raise ValueError(
            "density_to_stl: no surface found at threshold — check density field values")
    write_stl(output_path, rec.cpu().numpy())
    return output_path
