"""GPU replacement for the direct sparse solve in newton.py:52 (scipy spsolve).

Solves H x = b with Jacobi-preconditioned CG (sk_pcg_csr) on the GPU, for
sparse or dense SPD H. Raises numpy.linalg.LinAlgError when CG detects that H
is not positive definite.
"""

import ctypes
import warnings

import numpy as np
import scipy.sparse as sps

from . import _lib
from ._device import I32, dev, empty, host, ptr, stream


def pcg_solve(H, b, rtol=1e-12, maxit=None, return_info=False, allocate=True):
    """allocate=False runs the library's allocation-free entry point (sk_pcg_csr_noalloc: a static
    device workspace, at most 524,288 unknowns); the answer is the same bit for bit."""
    H = sps.csr_matrix(H)
    n = H.shape[0]
    if H.shape[0] != H.shape[1]:
        raise ValueError("pcg_solve: H must be square")
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    if n == 0:
        return (np.zeros(0), {"iters": 0, "relres": 0.0}) if return_info else np.zeros(0)
    maxit = 10 * n + 100 if maxit is None else maxit
    ip = dev(H.indptr.astype(np.int32), I32)
    ix = dev(H.indices.astype(np.int32), I32)
    vals = dev(H.data.astype(np.float64))
    bd = dev(b)
    x = empty(n)
    iters = ctypes.c_int64()
    relres = ctypes.c_double()
    entry = _lib.load().sk_pcg_csr if allocate else _lib.load().sk_pcg_csr_noalloc
    status = entry(n, ptr(ip), ptr(ix), ptr(vals), ptr(bd), ptr(x), float(rtol), int(maxit), ctypes.byref(iters),
                   ctypes.byref(relres), stream())
    if status == _lib.SK_ERR_PCG_BREAKDOWN:
        raise np.linalg.LinAlgError("Hessian is not positive definite (CG breakdown); use psd=True")
    if status == _lib.SK_ERR_PCG_MAXIT:
        warnings.warn(f"pcg_solve: {iters.value} iterations reached before rtol={rtol} "
                      f"(relative residual {relres.value:.3e})", RuntimeWarning)
    elif status != _lib.SK_OK:
        _lib.check(status)
    out = host(x)
    info = {"iters": iters.value, "relres": relres.value}
    return (out, info) if return_info else out
