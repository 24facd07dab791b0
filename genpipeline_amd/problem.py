"""Device-resident SIMP problem: a thin owner of one ``gp_problem`` handle.

Tensors live in the PyTorch-ROCm caching allocator; the library is handed
their device pointers and the current HIP stream. Densities and sensitivities
are float64 (element fields are small); PCG vectors are float32 (bulk) or
float64, with FP64 dot products in both cases.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import _lib as L

LAWS = {"clamp": L.GP_LAW_CLAMP, "emin": L.GP_LAW_EMIN}
OCS = {"cpu": L.GP_OC_CPU, "gpu": L.GP_OC_GPU}
FILTER_WEIGHTS = {"f64": L.GP_FILTER_F64, "f32": L.GP_FILTER_F32}
# "fp16" / "bf16": 16-bit PCG vector storage with float arithmetic, only as the
# inner solve of FP64 iterative refinement (refine=True)
PRECISIONS = {"fp32": L.GP_F32, "fp64": L.GP_F64, "fp16": L.GP_F16, "bf16": L.GP_BF16}


@dataclass(frozen=True)
class ProblemSpec:
    """Grid, material and SIMP parameters of one problem (serialisable)."""
    nx: int
    ny: int
    nz: int
    penal: float = 3.0
    rmin: float = 1.5
    volfrac: float = 0.3
    move: float = 0.2
    E0: float = 1.0
    Emin: float = 1e-9
    nu: float = 0.3
    law: str = "emin"          # "emin": Emin + x^p (E0-Emin); "clamp": max(x,1e-3)^p
    oc: str = "cpu"            # OC constants of SIMPSolver ("cpu") or SIMPSolverGPU ("gpu")
    # filter weights rmin - |d| in float64 (SIMPSolver) or rounded to float32
    # (SIMPSolverGPU, simp_solver_gpu.py:915)
    filter_weights: str = "f64"

    @property
    def n_elem(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def n_nodes(self) -> int:
        return (self.nx + 1) * (self.ny + 1) * (self.nz + 1)

    @property
    def n_dof(self) -> int:
        return 3 * self.n_nodes

    def to_c(self, ke: Optional[np.ndarray] = None):
        s = L.GpSpec()
        s.nx, s.ny, s.nz = self.nx, self.ny, self.nz
        s.law = LAWS[self.law]
        s.oc_variant = OCS[self.oc]
        s.E0, s.Emin, s.nu, s.penal, s.rmin, s.move = (self.E0, self.Emin, self.nu, self.penal,
                                                       self.rmin, self.move)
        s.filter_weights = FILTER_WEIGHTS[self.filter_weights]
        keep = None
        if ke is None:
            s.ke_kind = L.GP_KE_HEX8
            s.ke = None
        else:
            keep = np.ascontiguousarray(ke, dtype=np.float64).reshape(24, 24)
            s.ke_kind = L.GP_KE_CUSTOM
            s.ke = keep.ctypes.data_as(C.POINTER(C.c_double))
        return s, keep


def footprint_bytes(spec: ProblemSpec, precision: str = "fp32", recycle: int = 0) -> int:
    """Device bytes one problem holds (used for admission by HBM footprint),
    with the ring of `recycle` stored solutions when it recycles starts."""
    cs, _keep = spec.to_c()
    out = C.c_size_t(0)
    L.call("gp_footprint_bytes", C.byref(cs), PRECISIONS[precision], C.byref(out))
    ring = C.c_size_t(0)
    if recycle:
        L.call("gp_recycle_bytes", C.byref(cs), int(recycle), C.byref(ring))
    return int(out.value) + int(ring.value)


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(device):
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


@dataclass
class PcgResult:
    status: int
    iters: int
    outer: int
    relres: float

    @property
    def status_name(self) -> str:
        return L.STATUS_NAMES.get(self.status, str(self.status))


@dataclass
class SimpResult:
    x: torch.Tensor
    xPhys: torch.Tensor
    u: torch.Tensor
    compliance: float
    iters_done: int
    pcg_iters: int
    bisect_steps: int
    status: int
    seconds: float
    extra: dict = field(default_factory=dict)


class Problem:
    """One SIMP problem on one GPU.

    fixed: int64 array of fixed DOFs; f: float64 load vector (n_dof);
    preserved: optional bool array (n_elem) of non-design elements;
    ke: optional custom 24x24 element table (default: exact hex8).
    """

    def __init__(self, spec: ProblemSpec, fixed, f, preserved=None, ke=None, device=None):
        if not torch.cuda.is_available():
            raise RuntimeError("genpipeline_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.spec = spec
        d = torch.device("cuda") if device is None else torch.device(device)
        self.device = d if d.index is not None else torch.device("cuda", torch.cuda.current_device())
        L.load()
        cs, self._ke_keep = spec.to_c(ke)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            L.call("gp_problem_create", C.byref(cs), C.byref(h))
        self._h = h
        self.set_bcs(fixed, f)
        self.set_preserved(preserved)
        if ke is None:
            self._apply_tuned_tiles()

    def _apply_tuned_tiles(self):
        """Use the autotuner's cached K.u tile for this grid when there is one."""
        from . import autotune
        g = (self.spec.nx, self.spec.ny, self.spec.nz)
        for prec in ("fp32", "fp64"):
            t = autotune.lookup(g, prec, self.device)
            if t is not None:
                self.set_ku_tile(*t, precision=prec)

    # ------------------------------------------------------------ lifecycle
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.load().gp_problem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # -------------------------------------------------------------- set-up
    def set_bcs(self, fixed, f):
        fixed = np.ascontiguousarray(np.asarray(fixed, dtype=np.int64).ravel())
        f = np.ascontiguousarray(np.asarray(f, dtype=np.float64).ravel())
        if f.size != self.spec.n_dof:
            raise ValueError(f"f has {f.size} entries, expected n_dof={self.spec.n_dof}")
        self.fixed = fixed
        self.f_host = f
        with torch.cuda.device(self.device):
            L.call("gp_problem_set_bcs", self._h, fixed.ctypes.data_as(C.c_void_p), fixed.size,
                   f.ctypes.data_as(C.c_void_p))
        self.f = torch.from_numpy(f).to(self.device)

    def set_preserved(self, preserved):
        self.preserved = None
        arg = None
        if preserved is not None:
            m = np.ascontiguousarray(np.asarray(preserved).ravel().astype(np.uint8))
            if m.size != self.spec.n_elem:
                raise ValueError("preserved mask must have n_elem entries")
            self.preserved = m.astype(bool)
            arg = m.ctypes.data_as(C.c_void_p)
        with torch.cuda.device(self.device):
            L.call("gp_problem_set_preserved", self._h, arg)

    def set_profiling(self, every: int):
        """Bracket every `every`-th chunk of PCG iteration launches with HIP
        events (0 = off)."""
        L.call("gp_problem_set_profiling", self._h, int(every))

    def ku_profile(self, reset=True):
        """(mean milliseconds per PCG iteration launch, launches timed) since
        the last reset."""
        ms = C.c_double(0)
        n = C.c_int64(0)
        L.call("gp_problem_get_profile", self._h, C.byref(ms), C.byref(n), 1 if reset else 0)
        return (ms.value / n.value if n.value else float("nan")), int(n.value)

    KU_TILES = [(8, 32), (16, 16), (14, 18), (6, 34), (14, 34), (4, 64), (10, 26), (7, 34)]

    def set_ku_tile(self, tj: int, tk: int = 0, ni: int = 1, precision: str = "fp32"):
        """Modal K.u workgroup shape (tj x tk element columns, ni planes);
        tj = 0 selects the dense gather kernel."""
        L.call("gp_problem_set_ku_tile", self._h, PRECISIONS[precision], int(tj), int(tk), int(ni))

    def ku_tile(self, precision: str = "fp32"):
        """(tj, tk, ni, modal_ok) currently used for K.u."""
        o = [C.c_int32() for _ in range(4)]
        L.call("gp_problem_get_ku_tile", self._h, PRECISIONS[precision], *[C.byref(x) for x in o])
        return tuple(x.value for x in o)

    # ------------------------------------------------------------ helpers
    def _f64(self, a, n):
        t = torch.as_tensor(a, dtype=torch.float64, device=self.device).contiguous()
        if t.numel() != n:
            raise ValueError(f"expected {n} entries, got {t.numel()}")
        return t

    def stiffness(self, xPhys) -> torch.Tensor:
        x = self._f64(xPhys, self.spec.n_elem)
        E = torch.empty(self.spec.n_elem, dtype=torch.float32, device=self.device)
        L.call("gp_stiffness", self._h, _ptr(x), _ptr(E), _stream(self.device))
        return E

    def ku(self, E, u, mask=True) -> torch.Tensor:
        """y = K(E) u on the structured grid (fixed DOFs zeroed when mask)."""
        E = torch.as_tensor(E, device=self.device).to(torch.float32).contiguous()
        u = torch.as_tensor(u, device=self.device).contiguous()
        if u.dtype not in (torch.float32, torch.float64):
            u = u.to(torch.float64)
        if u.numel() != self.spec.n_dof or E.numel() != self.spec.n_elem:
            raise ValueError("shape mismatch in ku")
        y = torch.empty_like(u)
        dt = L.GP_F32 if u.dtype == torch.float32 else L.GP_F64
        L.call("gp_ku", self._h, _ptr(E), _ptr(u), _ptr(y), dt, 1 if mask else 0, _stream(self.device))
        return y

    def diag(self, E, dtype=torch.float64) -> torch.Tensor:
        E = torch.as_tensor(E, device=self.device).to(torch.float32).contiguous()
        d = torch.empty(self.spec.n_dof, dtype=dtype, device=self.device)
        L.call("gp_diag", self._h, _ptr(E), _ptr(d), L.GP_F32 if dtype == torch.float32 else L.GP_F64,
               _stream(self.device))
        return d

    def pcg(self, E, f=None, u0=None, tol=1e-6, max_iter=10000, precision="fp32", refine=False,
            inner_tol=1e-3, max_outer=50):
        """Jacobi-PCG on the free DOFs. Returns (u, PcgResult). u is float64
        when refine or precision == "fp64", float32 otherwise."""
        E = torch.as_tensor(E, device=self.device).to(torch.float32).contiguous()
        f = self.f if f is None else self._f64(f, self.spec.n_dof)
        wide = refine or precision == "fp64"
        dt = torch.float64 if wide else torch.float32
        if u0 is None:
            u = torch.zeros(self.spec.n_dof, dtype=dt, device=self.device)
        else:
            u = torch.as_tensor(u0, device=self.device).to(dt).contiguous().clone()
        o = L.GpPcgOpts()
        o.precision = PRECISIONS[precision]
        o.max_iter = int(max_iter)
        o.tol = float(tol)
        o.use_x0 = 0 if u0 is None else 1
        o.refine = 1 if refine else 0
        o.inner_tol = float(inner_tol)
        o.max_outer = int(max_outer)
        st = L.GpPcgStats()
        L.call("gp_pcg", self._h, _ptr(E), _ptr(f), _ptr(u), C.byref(o), C.byref(st), _stream(self.device))
        return u, PcgResult(st.status, st.iters, st.outer, st.relres)

    def recycle_guess(self, E, V):
        """Galerkin start u0 = V c, (V^T K V) c = V^T f, over the float32
        vectors V (1..8, n_dof each) at moduli E (gp_recycle_guess). Returns
        (u0 float32, c as a list of len(V) floats, number of vectors kept)."""
        E = torch.as_tensor(E, device=self.device).to(torch.float32).contiguous()
        V = [torch.as_tensor(v, device=self.device).to(torch.float32).contiguous() for v in V]
        if any(v.numel() != self.spec.n_dof for v in V):
            raise ValueError("every vector needs n_dof entries")
        ptrs = (C.c_void_p * len(V))(*[v.data_ptr() for v in V])
        u0 = torch.empty(self.spec.n_dof, dtype=torch.float32, device=self.device)
        coef = (C.c_double * 9)()
        L.call("gp_recycle_guess", self._h, _ptr(E), ptrs, len(V), _ptr(u0), coef,
               _stream(self.device))
        return u0, [coef[i] for i in range(len(V))], int(coef[8])

    def direct(self, xPhys, f=None):
        """Exact FP64 solve K(xPhys) u = f (banded LU with partial pivoting,
        gp_direct); returns (u float64, PcgResult with status only)."""
        x = self._f64(xPhys, self.spec.n_elem)
        f = self.f if f is None else self._f64(f, self.spec.n_dof)
        u = torch.empty(self.spec.n_dof, dtype=torch.float64, device=self.device)
        st = L.GpPcgStats()
        L.call("gp_direct", self._h, _ptr(x), _ptr(f), _ptr(u), C.byref(st), _stream(self.device))
        return u, PcgResult(st.status, st.iters, st.outer, st.relres)

    def sens(self, xPhys, u, with_energy=False):
        """dc = -dE/dx * u_e^T KE u_e (float64). With with_energy also returns
        sum_e E_e u_e^T KE u_e."""
        x = self._f64(xPhys, self.spec.n_elem)
        u = torch.as_tensor(u, device=self.device).contiguous()
        dt = L.GP_F32 if u.dtype == torch.float32 else L.GP_F64
        if dt == L.GP_F64 and u.dtype != torch.float64:
            u = u.to(torch.float64)
        dc = torch.empty(self.spec.n_elem, dtype=torch.float64, device=self.device)
        ce = torch.zeros(1, dtype=torch.float64, device=self.device) if with_energy else None
        L.call("gp_sens", self._h, _ptr(x), _ptr(u), dt, _ptr(dc), _ptr(ce), _stream(self.device))
        return (dc, float(ce.item())) if with_energy else dc

    def filter(self, v, mode=0) -> torch.Tensor:
        """mode 0: (H v)/Hs ; mode 1: H^T (v/Hs)."""
        v = self._f64(v, self.spec.n_elem)
        out = torch.empty_like(v)
        L.call("gp_filter", self._h, _ptr(v), _ptr(out), int(mode), _stream(self.device))
        return out

    def oc(self, x, dc, volfrac=None):
        """OC bisection; returns (xnew, xPhys_new, lmid, steps)."""
        x = self._f64(x, self.spec.n_elem)
        dc = self._f64(dc, self.spec.n_elem)
        xn = torch.empty_like(x)
        xp = torch.empty_like(x)
        st = L.GpOcStats()
        vf = self.spec.volfrac if volfrac is None else volfrac
        L.call("gp_oc", self._h, _ptr(x), _ptr(dc), float(vf), _ptr(xn), _ptr(xp), C.byref(st),
               _stream(self.device))
        return xn, xp, st.lmid, st.steps

    def compliance(self, u) -> float:
        u = torch.as_tensor(u, device=self.device).contiguous()
        dt = L.GP_F32 if u.dtype == torch.float32 else L.GP_F64
        out = C.c_double(0.0)
        L.call("gp_compliance", self._h, _ptr(u), dt, C.byref(out), _stream(self.device))
        return out.value

    def run(self, n_iters, x0=None, volfrac=None, tol=1e-6, max_iter=10000, precision="fp32",
            refine=False, inner_tol=1e-3, max_outer=50, warm_start=True, tol_schedule=False,
            early_stop=False, compliance_each_iter=False, u0=None, xphys0=None,
            warm_compliance=False, direct=False, recycle=0) -> SimpResult:
        """Whole SIMP loop on the device (gp_simp_run). warm_compliance: the
        compliance solve after each update starts from the previous u
        instead of zero. direct: every solve is the exact FP64 banded LU
        (the PCG arguments are then ignored). recycle = m (1..8, any PCG
        mode): each warm-started solve starts from the Galerkin combination of
        the run's last m solutions (float copies) instead of the last one
        alone."""
        sp = self.spec
        vf = sp.volfrac if volfrac is None else volfrac
        if x0 is None:
            x = torch.full((sp.n_elem,), float(vf), dtype=torch.float64, device=self.device)
            if self.preserved is not None:
                x[torch.from_numpy(self.preserved).to(self.device)] = 1.0
        else:
            x = self._f64(x0, sp.n_elem).clone()
        xPhys = x.clone() if xphys0 is None else self._f64(xphys0, sp.n_elem).clone()
        u = (torch.zeros(sp.n_dof, dtype=torch.float64, device=self.device) if u0 is None
             else self._f64(u0, sp.n_dof).clone())
        o = L.GpSimpOpts()
        o.n_iters = int(n_iters)
        o.volfrac = float(vf)
        o.pcg.precision = PRECISIONS[precision]
        o.pcg.max_iter = int(max_iter)
        o.pcg.tol = float(tol)
        o.pcg.use_x0 = 0
        o.pcg.refine = 1 if refine else 0
        o.pcg.inner_tol = float(inner_tol)
        o.pcg.max_outer = int(max_outer)
        o.tol_schedule = 1 if tol_schedule else 0
        o.early_stop = 1 if early_stop else 0
        o.compliance_each_iter = 1 if compliance_each_iter else 0
        o.warm_start = 1 if warm_start else 0
        o.warm_compliance = 1 if warm_compliance else 0
        o.direct = 1 if direct else 0
        o.recycle = int(recycle)
        st = L.GpSimpStats()
        L.call("gp_simp_run", self._h, C.byref(o), _ptr(x), _ptr(xPhys), _ptr(u), C.byref(st),
               _stream(self.device))
        return SimpResult(x, xPhys, u, st.compliance, st.iters_done, st.pcg_iters_total,
                          st.bisect_steps_total, st.status, st.seconds)
