"""Kinetic (inertial) terms on the GPU (simkit/energies/kinetic.py).

``0.5 c / h^2 (x - x~)^T M (x - x~)``: backward Euler with
x~ = x_curr + h (x_curr - x_prev)/h and c = 1 (:41-194, :291-345), BDF2 with
x~ = (4/3) x_curr - (1/3) x_prev + (8h/9) v_curr - (2h/9) v_prev and c = 9/4
(:63-101, :200-279). M may be any scipy sparse matrix (the lumped diagonal
mass in practice); the product M d runs as a CSR SpMV.
"""

import ctypes

import numpy as np
import scipy.sparse as sps

from .. import _lib
from .._device import I32, dev, empty, host, ptr, stream

_BE_COEFF = 1.0
_BDF2_COEFF = 9.0 / 4.0


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


def _run_bdf2(shape, x, x_curr, x_prev, x_prev2, x_prev3, M=None, h=1.0, energy=False, grad=False, target=False,
              velocity=False):
    n = int(np.prod(shape))
    arrs = [None if a is None else dev(np.asarray(a, dtype=np.float64).reshape(-1))
            for a in (x, x_curr, x_prev, x_prev2, x_prev3)]
    ip = ix = vals = None
    if M is not None:
        ip, ix, vals = _csr(M)
    e = ctypes.c_double()
    g = empty(n) if grad else None
    xt = empty(n) if target else None
    v = empty(n) if velocity else None
    _lib.call("sk_kinetic_bdf2", n, *[ptr(a) for a in arrs], float(h), ptr(ip), ptr(ix), ptr(vals),
              ctypes.byref(e) if energy else None, ptr(g), ptr(xt), ptr(v), stream())
    return e.value, g, xt, v


def velocity_bdf2(x_curr, x_prev, x_prev2, h):
    """(3 x_curr - 4 x_prev + x_prev2) / (2 h)."""
    shape = np.shape(x_curr)
    return host(_run_bdf2(shape, None, x_curr, x_prev, x_prev2, x_prev2, h=h, velocity=True)[3]).reshape(shape)


def bdf2_target(x_curr, x_prev, x_prev2, x_prev3, h):
    """(4/3) x_curr - (1/3) x_prev + (8h/9) v_curr - (2h/9) v_prev."""
    shape = np.shape(x_curr)
    return host(_run_bdf2(shape, None, x_curr, x_prev, x_prev2, x_prev3, h=h, target=True)[2]).reshape(shape)


def kinetic_energy_bdf2(x, x_curr, x_prev, x_prev2, x_prev3, M, h):
    return _run_bdf2(np.shape(x), x, x_curr, x_prev, x_prev2, x_prev3, M=M, h=h, energy=True)[0]


def kinetic_gradient_bdf2(x, x_curr, x_prev, x_prev2, x_prev3, M, h):
    shape = np.shape(x)
    return host(_run_bdf2(shape, x, x_curr, x_prev, x_prev2, x_prev3, M=M, h=h, grad=True)[1]).reshape(shape)


def kinetic_hessian_bdf2(M, h):
    return kinetic_hessian(M, h, _BDF2_COEFF)
