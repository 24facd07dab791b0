"""ctypes binding of the oracle's multi-core C restatement (simp_cpu.c).

TEST / BASELINE INFRASTRUCTURE ONLY (tests/ and bench.py's cpu_baseline).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_SO = Path(__file__).resolve().parent / "_build" / "libsimp_cpu.so"
_lib = None
_dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not _SO.exists():
            from .build_oracle import build
            build()
        L = C.CDLL(str(_SO))
        L.ocpu_ku.argtypes = [C.c_int] * 3 + [_dp, _dp, _dp, _dp, C.c_void_p]
        L.ocpu_ku.restype = None
        L.ocpu_pcg.argtypes = [C.c_int] * 3 + [_dp, _dp, _dp, _u8, _dp, C.c_int, C.c_double,
                                               C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.ocpu_pcg.restype = C.c_int
        L.ocpu_simp_run.argtypes = ([C.c_int] * 3 + [_dp, C.c_int] + [C.c_double] * 6 +
                                    [_dp, _u8, C.c_int, C.c_double, C.c_int, _dp, _dp, _dp,
                                     C.POINTER(C.c_double), C.c_int])
        L.ocpu_simp_run.restype = C.c_long
        L.ocpu_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _mask(fixed, n_dof):
    m = np.zeros(n_dof, np.uint8)
    m[np.asarray(fixed, dtype=np.int64)] = 1
    return m


def ku(grid, KE, E, u, fixed=None):
    nx, ny, nz = grid
    y = np.empty_like(u, dtype=np.float64)
    fm = None if fixed is None else _mask(fixed, u.size).ctypes.data_as(C.c_void_p)
    m = None if fixed is None else _mask(fixed, u.size)
    fm = None if m is None else m.ctypes.data_as(C.c_void_p)
    lib().ocpu_ku(nx, ny, nz, np.ascontiguousarray(KE, np.float64), np.ascontiguousarray(E, np.float64),
                  np.ascontiguousarray(u, np.float64), y, fm)
    return y


def pcg(grid, KE, E, f, fixed, tol=1e-8, max_iter=100000, u0=None):
    nx, ny, nz = grid
    u = np.zeros_like(f) if u0 is None else np.array(u0, np.float64)
    st = C.c_int(0)
    rr = C.c_double(0)
    it = lib().ocpu_pcg(nx, ny, nz, np.ascontiguousarray(KE, np.float64), np.ascontiguousarray(E, np.float64),
                        np.ascontiguousarray(f, np.float64), _mask(fixed, f.size), u,
                        0 if u0 is None else 1, tol, max_iter, C.byref(st), C.byref(rr))
    return u, it, st.value, rr.value


def simp_run(grid, KE, f, fixed, n_iters, volfrac, tol, law="emin", penal=3.0, E0=1.0, Emin=1e-9,
             rmin=1.5, move=0.2, max_iter=100000, x0=None, xphys0=None, u0=None, threads=0):
    """Returns (x, xPhys, u, compliance, total_pcg_iters)."""
    nx, ny, nz = grid
    ne = nx * ny * nz
    x = np.full(ne, float(volfrac)) if x0 is None else np.array(x0, np.float64)
    xp = x.copy() if xphys0 is None else np.array(xphys0, np.float64)
    u = np.zeros(f.size) if u0 is None else np.array(u0, np.float64)
    comp = C.c_double(0)
    its = lib().ocpu_simp_run(nx, ny, nz, np.ascontiguousarray(KE, np.float64), 0 if law == "clamp" else 1,
                              penal, E0, Emin, rmin, volfrac, move, np.ascontiguousarray(f, np.float64),
                              _mask(fixed, f.size), n_iters, tol, max_iter, x, xp, u, C.byref(comp),
                              threads)
    return x, xp, u, comp.value, int(its)


def max_threads() -> int:
    return int(lib().ocpu_max_threads())
