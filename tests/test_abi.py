"""The C-ABI library loads and exports every symbol include/gp_simp.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import re
from pathlib import Path

from genpipeline_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "gp_simp.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\*|int)\s+(gp_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_symbols():
    names = declared_symbols()
    assert "gp_pcg" in names and "gp_ku" in names and "gp_simp_run" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.lib_path()))
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_lib.exported_symbols()) == declared_symbols()


def test_version_and_error_string():
    lib = _lib.load()
    assert lib.gp_version() >= 100
    assert isinstance(lib.gp_last_error(), bytes)


def test_argument_errors_without_gpu():
    """Argument validation happens before any device call."""
    lib = _lib.load()
    spec = _lib.GpSpec()
    spec.nx, spec.ny, spec.nz = 0, 4, 4
    spec.rmin = 1.5
    spec.nu = 0.3
    out = ctypes.c_size_t()
    assert lib.gp_footprint_bytes(ctypes.byref(spec), 0, ctypes.byref(out)) == _lib.GP_ERR_ARG
    assert b"grid" in lib.gp_last_error()
    spec.nx = 4
    spec.rmin = 3.5
    assert lib.gp_footprint_bytes(ctypes.byref(spec), 0, ctypes.byref(out)) == _lib.GP_ERR_ARG
    spec.rmin = 1.5
    assert lib.gp_footprint_bytes(ctypes.byref(spec), 0, ctypes.byref(out)) == _lib.GP_OK
    assert out.value > 0
