"""The allocation-free generic PCG (sk_pcg_csr_noalloc, a static device
workspace) against the allocating entry point it mirrors (sk_pcg_csr), through
the C ABI: the same kernels and launch shapes, so results must be bitwise
identical. Two calls in a row, the second on a smaller system (the workspace
keeps the first call's contents), and the size limit as an error."""

import ctypes

import numpy as np
import pytest

from oracle import solver
from oracle.mesh import unit_cube

pytestmark = pytest.mark.gpu


def _system(cells, seed):
    X, T = unit_cube(cells)
    p = solver.BEProblem(X, T, 3e5, 2e6, 1000.0, 1 / 60, np.where(X[:, 1] == 0)[0], gravity=-9.8)
    x = X.reshape(-1) + 0.02 * np.random.default_rng(seed).standard_normal(X.size) / cells
    H = p.potential_hessian(x).tocsr()
    H.sort_indices()
    b = np.random.default_rng(seed + 1).standard_normal(H.shape[0])
    return H, b


def _solve(entry, H, b, rtol=1e-10, maxit=5000):
    import torch
    from simkit_amd import _lib
    from simkit_amd._device import I32, dev, empty, host, ptr, stream
    lib = _lib.load()
    n = H.shape[0]
    ip, ix = dev(H.indptr.astype(np.int32), I32), dev(H.indices.astype(np.int32), I32)
    vals, bd = dev(H.data.astype(np.float64)), dev(b)
    x = empty(n)
    x.fill_(float("nan"))                   # every entry must be written
    iters, relres = ctypes.c_int64(-1), ctypes.c_double(-1.0)
    fn = getattr(lib, entry)
    fn.argtypes = [ctypes.c_int64] + [ctypes.c_void_p] * 5 + [ctypes.c_double, ctypes.c_int64,
                                                              ctypes.POINTER(ctypes.c_int64),
                                                              ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
    status = fn(n, ptr(ip), ptr(ix), ptr(vals), ptr(bd), ptr(x), rtol, maxit, ctypes.byref(iters),
                ctypes.byref(relres), stream())
    torch.cuda.synchronize()
    return status, host(x), iters.value, relres.value


def test_noalloc_pcg_matches_allocating_entry_point():
    from simkit_amd import _lib
    for cells, seed in ((9, 0), (5, 3)):     # the second system is smaller than the first
        H, b = _system(cells, seed)
        print(f"noalloc: {cells}^3 cells, n = {H.shape[0]}", flush=True)
        s1, x1, it1, rr1 = _solve("sk_pcg_csr", H, b)
        s2, x2, it2, rr2 = _solve("sk_pcg_csr_noalloc", H, b)
        assert s1 == s2 == _lib.SK_OK
        assert it1 == it2 > 0 and rr1 == rr2 <= 1e-10
        assert np.array_equal(x1, x2)
        assert np.linalg.norm(H @ x2 - b) <= 1e-9 * np.linalg.norm(b)
    # an iteration limit hit inside the static workspace reports the same status and iterate
    s1, x1, it1, _ = _solve("sk_pcg_csr", H, b, maxit=7)
    s2, x2, it2, _ = _solve("sk_pcg_csr_noalloc", H, b, maxit=7)
    assert s1 == s2 == _lib.SK_ERR_PCG_MAXIT and it1 == it2 == 7 and np.array_equal(x1, x2)
    # the Python drop-in's switch
    from simkit_amd.linsolve import pcg_solve
    assert np.array_equal(pcg_solve(H, b, rtol=1e-10, allocate=False), pcg_solve(H, b, rtol=1e-10))


def test_noalloc_pcg_rejects_sizes_beyond_its_workspace():
    import torch
    from simkit_amd import _lib
    from simkit_amd._device import ptr, stream
    lib = _lib.load()
    buf = torch.zeros(16, dtype=torch.float64, device="cuda")
    ibuf = torch.zeros(16, dtype=torch.int32, device="cuda")
    fn = lib.sk_pcg_csr_noalloc
    fn.argtypes = [ctypes.c_int64] + [ctypes.c_void_p] * 5 + [ctypes.c_double, ctypes.c_int64,
                                                              ctypes.POINTER(ctypes.c_int64),
                                                              ctypes.POINTER(ctypes.c_double), ctypes.c_void_p]
    it, rr = ctypes.c_int64(), ctypes.c_double()
    # the size check comes before any device access, so the small buffers are never read
    status = fn(524289, ptr(ibuf), ptr(ibuf), ptr(buf), ptr(buf), ptr(buf), 1e-8, 10, ctypes.byref(it),
                ctypes.byref(rr), stream())
    assert status == _lib.SK_ERR_INVALID
    assert b"static workspace" in lib.sk_last_error()
