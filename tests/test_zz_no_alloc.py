"""gp_csr_pcg_noalloc (CSR Jacobi-PCG whose workspace is the library's
static device array) against gp_csr_pcg (library-allocated workspace) through
the C ABI: bitwise the same iterate and statistics on two calls in a row,
the second on a smaller system, and a clear error beyond the size limit."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp
import torch

pytestmark = pytest.mark.gpu

N_MAX = 1 << 20


def _system(n_side, seed):
    """A 2-D Laplacian-plus-shift on an n_side^2 grid with a few fixed rows."""
    rng = np.random.default_rng(seed)
    m = n_side
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(m, m))
    A = (sp.kronsum(T, T) + 0.01 * sp.identity(m * m)).tocsr()
    n = A.shape[0]
    b = rng.standard_normal(n).astype(np.float32)
    fixed = np.sort(rng.choice(n, 5, replace=False)).astype(np.int64)
    return A, b, fixed


def _call(name, A, b, fixed, dev, max_iter=400, tol=1e-6):
    from genpipeline_amd import _lib as L
    rp = torch.from_numpy(A.indptr.astype(np.int64)).to(dev)
    ci = torch.from_numpy(A.indices.astype(np.int64)).to(dev)
    v = torch.from_numpy(A.data.astype(np.float32)).to(dev)
    d = torch.from_numpy(A.diagonal().astype(np.float32)).to(dev)
    bb = torch.from_numpy(b).to(dev)
    x = torch.zeros_like(bb)
    fx = torch.from_numpy(fixed).to(dev)
    st = L.GpPcgStats()
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    L.call(name, C.c_void_p(rp.data_ptr()), C.c_void_p(ci.data_ptr()), C.c_void_p(v.data_ptr()),
           C.c_void_p(d.data_ptr()), C.c_void_p(bb.data_ptr()), C.c_void_p(x.data_ptr()),
           C.c_void_p(fx.data_ptr()), fx.numel(), bb.numel(), max_iter, tol, C.byref(st), stream)
    torch.cuda.synchronize(dev)
    return x.cpu().numpy(), (st.status, st.iters, st.relres)


def test_noalloc_equals_allocating_solve_twice(gpu):
    for n_side, seed in ((300, 1), (120, 2)):  # 90,000 then 14,400 unknowns
        A, b, fixed = _system(n_side, seed)
        print(f"[noalloc] n = {A.shape[0]}", flush=True)
        xa, sa = _call("gp_csr_pcg", A, b, fixed, gpu)
        xn, sn = _call("gp_csr_pcg_noalloc", A, b, fixed, gpu)
        assert sa == sn, (sa, sn)
        assert np.array_equal(xa, xn)
        assert sa[0] == 0 and np.all(xn[fixed] == 0)
        # and it solves the system on the free rows (FP32 vectors: the true
        # residual sits a few times above the recursive one's 1e-6)
        free = np.setdiff1d(np.arange(A.shape[0]), fixed)
        r = b[free] - (A @ xn.astype(np.float64))[free]
        assert np.linalg.norm(r) <= 1e-5 * np.linalg.norm(b)


def test_noalloc_rejects_systems_beyond_its_array(gpu):
    from genpipeline_amd import _lib as L
    one = torch.zeros(1, dtype=torch.float32, device=gpu)
    idx = torch.zeros(2, dtype=torch.int64, device=gpu)
    st = L.GpPcgStats()
    with pytest.raises(L.GpError, match="no-allocation limit"):
        L.call("gp_csr_pcg_noalloc", C.c_void_p(idx.data_ptr()), C.c_void_p(idx.data_ptr()),
               C.c_void_p(one.data_ptr()), C.c_void_p(one.data_ptr()), C.c_void_p(one.data_ptr()),
               C.c_void_p(one.data_ptr()), None, 0, N_MAX + 1, 10, 1e-6, C.byref(st), None)
