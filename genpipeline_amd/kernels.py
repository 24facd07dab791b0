"""Explicit-table kernel functions with the signatures of
genpipeline/cuda_kernels/__init__.py:286-502.

These take the element matrix and the edof table as tensors, as the
reference's kernels do, and run the library's generic gather kernels
(``gp_generic_sens``, ``gp_generic_spmv``) and CSR Jacobi-PCG
(``gp_csr_pcg``). They have no Python fallback: without the HIP library or
a GPU they raise. Structured-grid callers use ``Problem`` instead, which
derives element connectivity arithmetically.

Float64 inputs run in float64, as the reference kernels read double; float32
inputs in float32. Element products are formed in FP64, and results come back
in the dtype of the input they mirror.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L


def _dev(t: torch.Tensor, dtype) -> torch.Tensor:
    if not t.is_cuda:
        raise RuntimeError("genpipeline_amd kernels take device tensors")
    return t.to(dtype).contiguous()


def _stream(t: torch.Tensor):
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _ftype(*ts):
    """float64 when any float input is float64 (the reference kernels' type),
    else float32."""
    return torch.float64 if any(t.dtype == torch.float64 for t in ts) else torch.float32


def _edof(edof_mat: torch.Tensor, n_dof: int) -> torch.Tensor:
    ed = _dev(edof_mat, torch.int32)
    if ed.numel() and (int(ed.min()) < 0 or int(ed.max()) >= n_dof):
        raise ValueError("edof_mat entries must lie in [0, n_dof)")
    return ed


def simp_sensitivity(xPhys: torch.Tensor, u: torch.Tensor, Ke: torch.Tensor,
                     edof_mat: torch.Tensor, penal: float, nx: int, ny: int, nz: int) -> torch.Tensor:
    """dc_e = -penal max(x_e, 1e-3)^(penal-1) u_e^T Ke u_e, shaped like xPhys
    (__init__.py:286-326; kernel simp_sensitivity_ptx.cu:163-270)."""
    n = int(nx) * int(ny) * int(nz)
    ft = _ftype(xPhys, u, Ke)
    x = _dev(xPhys, ft)
    if x.numel() != n or edof_mat.shape[0] != n:
        raise ValueError("xPhys and edof_mat must hold nx*ny*nz elements")
    uu, ke = _dev(u, ft), _dev(Ke, ft)
    ed = _edof(edof_mat, uu.numel())
    dc = torch.empty_like(x)
    L.call("gp_generic_sens", C.c_void_p(x.data_ptr()), C.c_void_p(uu.data_ptr()),
           C.c_void_p(ke.data_ptr()), C.c_void_p(ed.data_ptr()), n, float(penal),
           L.GP_F64 if ft == torch.float64 else L.GP_F32, C.c_void_p(dc.data_ptr()), _stream(x))
    return dc.view(xPhys.shape).to(xPhys.dtype)


def simp_sensitivity_aggressive(xPhys, u, Ke, edof_mat, penal, nx, ny, nz) -> torch.Tensor:
    """Same result as simp_sensitivity (__init__.py:476-502)."""
    return simp_sensitivity(xPhys, u, Ke, edof_mat, penal, nx, ny, nz)


def fused_spmv(xPhys: torch.Tensor, p: torch.Tensor, Ke: torch.Tensor, edof_mat: torch.Tensor,
               penal: float, nx: int, ny: int, nz: int) -> torch.Tensor:
    """y = sum_e max(x_e, 1e-3)^penal Ke p_e scattered by edof, shaped like p
    (__init__.py:329-368; kernel simp_sensitivity_ptx.cu:301-373). Every DOF
    sums its terms in a fixed order, so the result does not vary run to run
    (the reference kernel adds them atomically)."""
    n = int(nx) * int(ny) * int(nz)
    ft = _ftype(xPhys, p, Ke)
    x = _dev(xPhys, ft)
    if x.numel() != n or edof_mat.shape[0] != n:
        raise ValueError("xPhys and edof_mat must hold nx*ny*nz elements")
    pp, ke = _dev(p, ft), _dev(Ke, ft)
    ed = _edof(edof_mat, pp.numel())
    nw = C.c_int64()
    L.call("gp_generic_spmv_workspace", n, pp.numel(), C.byref(nw))
    ws = torch.empty(int(nw.value), dtype=torch.int32, device=x.device)
    y = torch.empty_like(pp)
    L.call("gp_generic_spmv", C.c_void_p(x.data_ptr()), C.c_void_p(pp.data_ptr()),
           C.c_void_p(ke.data_ptr()), C.c_void_p(ed.data_ptr()), n, pp.numel(), float(penal),
           L.GP_F64 if ft == torch.float64 else L.GP_F32, C.c_void_p(ws.data_ptr()),
           C.c_void_p(y.data_ptr()), _stream(x))
    return y.view(p.shape).to(p.dtype)


def pcg_solve_cuda(row_ptr: torch.Tensor, col_idx: torch.Tensor, values: torch.Tensor,
                   diagonal: torch.Tensor, b: torch.Tensor, x0: torch.Tensor,
                   fixed_dofs: torch.Tensor, max_iter: int = 100, tol: float = 1e-6) -> torch.Tensor:
    """Jacobi-PCG on a CSR matrix with fixed DOFs held at zero
    (__init__.py:443-473; simp_pcg_cuda.cu:66-124): x starts at x0, the
    residual starts at b, z = r / diagonal, stop when ||r|| < tol ||b|| or
    |p^T A p| < 1e-30."""
    rp, ci = _dev(row_ptr, torch.int64), _dev(col_idx, torch.int64)
    v, d = _dev(values, torch.float32), _dev(diagonal, torch.float32)
    bb = _dev(b, torch.float32)
    x = _dev(x0, torch.float32).clone()
    fx = _dev(fixed_dofs, torch.int64)
    st = L.GpPcgStats()
    L.call("gp_csr_pcg", C.c_void_p(rp.data_ptr()), C.c_void_p(ci.data_ptr()),
           C.c_void_p(v.data_ptr()), C.c_void_p(d.data_ptr()), C.c_void_p(bb.data_ptr()),
           C.c_void_p(x.data_ptr()), C.c_void_p(fx.data_ptr()) if fx.numel() else None,
           fx.numel(), bb.numel(), int(max_iter), float(tol), C.byref(st), _stream(bb))
    return x.to(x0.dtype)
