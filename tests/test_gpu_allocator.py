"""The library's device workspaces come from PyTorch's caching allocator
(gp_set_allocator, registered when the binding loads): a problem's buffers
count in torch.cuda.memory_allocated and go back when it closes, and the
allocator cannot change while library allocations are alive."""
import numpy as np
import pytest
import torch

from genpipeline_amd import Problem, ProblemSpec, footprint_bytes
from genpipeline_amd import _lib as L
from genpipeline_amd.bcs import clamp_x0
from genpipeline_amd.problems import point_load_vector

pytestmark = pytest.mark.gpu


def _problem(dev, nx=16, ny=8, nz=8):
    f = point_load_vector(nx, ny, nz, np.random.default_rng(0), 2)
    return Problem(ProblemSpec(nx, ny, nz), clamp_x0(nx, ny, nz), f, device=dev)


def test_workspaces_are_torch_allocations(gpu):
    if not L.torch_allocator_enabled():
        pytest.skip("GP_TORCH_ALLOCATOR=0: the library uses hipMalloc")
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated(gpu)
    P = _problem(gpu)
    torch.cuda.synchronize()
    during = torch.cuda.memory_allocated(gpu)
    assert during > before
    # a solve allocates its PCG workspace lazily, through the same allocator;
    # the footprint counts the set-up and that workspace
    x = torch.full((P.spec.n_elem,), 0.3, dtype=torch.float64, device=gpu)
    u, r = P.pcg(P.stiffness(x), tol=1e-6, max_iter=2000)
    assert r.status == 0
    torch.cuda.synchronize()
    mid = torch.cuda.memory_allocated(gpu)
    fp = footprint_bytes(P.spec)
    assert mid - before >= 0.5 * fp, (before, during, mid, fp)
    with pytest.raises(L.GpError) as e:
        L.call("gp_set_allocator", L.GP_ALLOC_FN(), L.GP_FREE_FN(), None)
    assert e.value.code == L.GP_ERR_STATE
    P.close()
    torch.cuda.synchronize()
    # the library's buffers went back to the allocator
    assert mid - torch.cuda.memory_allocated(gpu) >= 0.5 * fp


def test_results_unchanged_by_the_allocator(gpu):
    """The same solve twice (buffers reused from the cache the second time)
    gives the same bits."""
    out = []
    for _ in range(2):
        P = _problem(gpu, 12, 6, 6)
        res = P.run(3, tol=1e-6, max_iter=5000, precision="fp32", warm_start=True)
        out.append(res)
        P.close()
    assert torch.equal(out[0].xPhys, out[1].xPhys) and out[0].compliance == out[1].compliance


_CHILD = r"""
import sys, numpy as np, torch
sys.path.insert(0, sys.argv[1])
from genpipeline_amd import Problem, ProblemSpec
from genpipeline_amd import _lib as L
from genpipeline_amd.bcs import clamp_x0
from genpipeline_amd.problems import point_load_vector
assert not L.torch_allocator_enabled()
nx, ny, nz = 12, 6, 6
f = point_load_vector(nx, ny, nz, np.random.default_rng(0), 2)
P = Problem(ProblemSpec(nx, ny, nz), clamp_x0(nx, ny, nz), f, device=torch.device("cuda", 0))
r = P.run(3, tol=1e-6, max_iter=5000, precision="fp32", warm_start=True, recycle=2)
np.save(sys.argv[2], r.xPhys.cpu().numpy())
print(float(r.compliance).hex())
P.close()
"""


def test_hipmalloc_path_gives_the_same_bits(gpu, tmp_path):
    """GP_TORCH_ALLOCATOR=0 keeps hipMalloc (in a child process, since the
    allocator is registered when the binding loads): the same run, with
    recycled starts, gives the same bits on both allocator paths."""
    import os
    import subprocess
    import sys
    from pathlib import Path
    root = str(Path(__file__).resolve().parents[1])
    out = tmp_path / "x.npy"
    env = dict(os.environ, GP_TORCH_ALLOCATOR="0")
    p = subprocess.run([sys.executable, "-c", _CHILD, root, str(out)], env=env,
                       capture_output=True, text=True, timeout=90)
    assert p.returncode == 0, p.stderr[-2000:]
    P = _problem(gpu, 12, 6, 6)
    r = P.run(3, tol=1e-6, max_iter=5000, precision="fp32", warm_start=True, recycle=2)
    P.close()
    assert p.stdout.strip().splitlines()[-1] == float(r.compliance).hex()
    assert np.array_equal(np.load(out), r.xPhys.cpu().numpy())
