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


def test_size_limits_without_gpu():
    """Maximum sizes are rejected before any allocation: a grid above 2^29
    nodes, an axis past 2^20 (whose node count would overflow 32 bits
    first), isosurface grids without a cube, and unsymmetric pre-processing
    taps."""
    lib = _lib.load()
    spec = _lib.GpSpec()
    spec.rmin = 1.5
    spec.nu = 0.3
    out = ctypes.c_size_t()
    for dims in ((1024, 1024, 1024), (2 ** 31 - 1, 1, 1), (1, 2 ** 21, 1)):
        spec.nx, spec.ny, spec.nz = dims
        assert lib.gp_footprint_bytes(ctypes.byref(spec), 0, ctypes.byref(out)) == _lib.GP_ERR_ARG
        h = ctypes.c_void_p()
        assert lib.gp_problem_create(ctypes.byref(spec), ctypes.byref(h)) == _lib.GP_ERR_ARG
        assert not h.value
    spec.nx, spec.ny, spec.nz = 511, 511, 511  # 2^27 nodes: accepted
    assert lib.gp_footprint_bytes(ctypes.byref(spec), 0, ctypes.byref(out)) == _lib.GP_OK
    n = ctypes.c_int64()
    assert lib.gp_mc_workspace(1, 5, 5, ctypes.byref(n)) == _lib.GP_ERR_ARG
    assert lib.gp_mc_workspace(3, 3, 3, ctypes.byref(n)) == _lib.GP_OK and n.value == 2
    w = (ctypes.c_double * 5)(0.1, 0.2, 0.4, 0.2, 0.3)  # not symmetric
    buf = ctypes.c_void_p(1)
    assert lib.gp_voxel_prep(buf, 1, 4, 4, 4, 2, 2, 2, w, 2, ctypes.c_double(15.0), buf, buf,
                             None) == _lib.GP_ERR_ARG
    assert b"symmetric" in lib.gp_last_error()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: any other --gpus value exits with status 2
    and says so, before touching a device."""
    import subprocess
    import sys
    root = HEADER.parent.parent
    p = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], cwd=root,
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 2
    assert "one GPU only" in p.stdout + p.stderr
    assert p.stdout.strip() == "" or not p.stdout.lstrip().startswith("{")


def test_recycle_ring_bytes_without_gpu():
    """The recycled-start ring of m solutions: 2m float vectors of n_dof (to
    256 B) plus a fixed FP64 Gram area, whatever the grid; 0 without it;
    counts above 8 rejected."""
    from genpipeline_amd import ProblemSpec, footprint_bytes
    base = {}
    for nx, ny, nz in ((8, 4, 4), (16, 8, 8)):
        sp = ProblemSpec(nx, ny, nz)
        b0 = footprint_bytes(sp)
        assert footprint_bytes(sp, recycle=0) == b0
        ring = {m: footprint_bytes(sp, recycle=m) - b0 for m in (1, 6, 8)}
        for m, b in ring.items():
            vec = (2 * m * sp.n_dof * 4 + 255) // 256 * 256
            base[(nx, m)] = b - vec
            assert b > vec
        assert ring[1] < ring[6] < ring[8]
    # the Gram area does not depend on the grid or the count
    assert len(set(base.values())) == 1
    lib = _lib.load()
    cs, _keep = ProblemSpec(8, 4, 4).to_c()
    out = ctypes.c_size_t()
    assert lib.gp_recycle_bytes(ctypes.byref(cs), 9, ctypes.byref(out)) == _lib.GP_ERR_ARG
