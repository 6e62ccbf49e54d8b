"""Backward-Euler kinetic (inertial) terms on the GPU (simkit/energies/kinetic.py:41-194, :291-345).

``0.5 c / h^2 (x - x~)^T M (x - x~)`` with x~ = x_curr + h (x_curr - x_prev)/h
and c = 1. M may be any scipy sparse matrix (the lumped diagonal mass in
practice); the product M d runs as a CSR SpMV. The BDF2 variants are not on
this path.
"""

import ctypes

import numpy as np
import scipy.sparse as sps

from .. import _lib
from .._device import I32, dev, empty, host, ptr, stream

_BE_COEFF = 1.0


def _csr(M):
    C = sps.csr_matrix(M)
    return (dev(C.indptr.astype(np.int32), I32), dev(C.indices.astype(np.int32), I32),
            dev(C.data.astype(np.float64)))


def _run(shape, x=None, x_curr=None, x_prev=None, M=None, h=1.0, c=_BE_COEFF, energy=False, grad=False,
         target=False, velocity=False):
    n = int(np.prod(shape))
    xd = None if x is None else dev(np.asarray(x, dtype=np.float64).reshape(-1))
    xc = None if x_curr is None else dev(np.asarray(x_curr, dtype=np.float64).reshape(-1))
    xp = None if x_prev is None else dev(np.asarray(x_prev, dtype=np.float64).reshape(-1))
    ip = ix = vals = None
    if M is not None:
        ip, ix, vals = _csr(M)
    e = ctypes.c_double()
    g = empty(n) if grad else None
    xt = empty(n) if target else None
    v = empty(n) if velocity else None
    _lib.call("sk_kinetic", n, ptr(xd), ptr(xc), ptr(xp), float(h), float(c), ptr(ip), ptr(ix), ptr(vals),
              ctypes.byref(e) if energy else None, ptr(g), ptr(xt), ptr(v), stream())
    return e.value, g, xt, v


def velocity_be(x_curr, x_prev, h):
    """(x_curr - x_prev) / h."""
    shape = np.shape(x_curr)
    return host(_run(shape, x_curr=x_curr, x_prev=x_prev, h=h, velocity=True)[3]).reshape(shape)


def be_target(x_curr, x_prev, h):
    """x_curr + h * velocity_be(x_curr, x_prev, h)."""
    shape = np.shape(x_curr)
    return host(_run(shape, x_curr=x_curr, x_prev=x_prev, h=h, target=True)[2]).reshape(shape)


def kinetic_energy(d, M, h, c):
    """0.5 c d^T M d / h^2 as a Python float."""
    return _run(np.shape(d), x=d, M=M, h=h, c=c, energy=True)[0]


def kinetic_gradient(d, M, h, c):
    """c (M d) / h^2."""
    shape = np.shape(d)
    return host(_run(shape, x=d, M=M, h=h, c=c, grad=True)[1]).reshape(shape)


def kinetic_hessian(M, h, c):
    """c / h^2 * M."""
    return M * (c / (h ** 2))


def kinetic_energy_be(x, x_curr, x_prev, M, h):
    return _run(np.shape(x), x=x, x_curr=x_curr, x_prev=x_prev, M=M, h=h, energy=True)[0]


def kinetic_gradient_be(x, x_curr, x_prev, M, h):
    shape = np.shape(x)
    return host(_run(shape, x=x, x_curr=x_curr, x_prev=x_prev, M=M, h=h, grad=True)[1]).reshape(shape)


def kinetic_hessian_be(M, h):
    return kinetic_hessian(M, h, _BE_COEFF)


def kinetic_energy_be_u(u, x_curr, x_prev, M, h, x_bar):
    return kinetic_energy_be(x_bar + u, x_curr, x_prev, M, h)


def kinetic_gradient_be_u(u, x_curr, x_prev, M, h, x_bar):
    return kinetic_gradient_be(x_bar + u, x_curr, x_prev, M, h)
