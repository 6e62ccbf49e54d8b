"""Oracle linear modal analysis (test infrastructure).

Restates linear_modal_analysis.py:18-73 with eigs.py:66-110: linear-elasticity
stiffness with mu = lam = 1 (no PSD projection), lumped mass with rho = 1,
constrained DOFs eliminated, and the k eigenpairs of (K + 1e-6 M) phi =
lambda M phi nearest zero by shift-invert Lanczos (scipy eigsh, sigma=0,
v0 = ones). The reference factors with UMFPACK (cvxopt, pinned nowhere and
absent here); scipy's default SuperLU factorization is used instead, which
changes only round-off. Pinned in the tests against a dense generalized
eigh on small meshes. Only the constrained (bI given) form is restated: the
unconstrained stiffness is singular and the shift-invert solve at sigma = 0
is undefined there.
"""

import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from . import fem


def stiffness_and_mass(X, T):
    D = fem.tet_operator(X, T)
    J = fem.deformation_jacobian(D, T, X.shape[0])
    vol = fem.tet_volumes(X, T)
    ones = np.ones(T.shape[0])
    K = fem.hessian_x("linear", X.reshape(-1), J, ones, ones, vol, psd=False)
    m = np.repeat(fem.lumped_mass(X, T, 1.0), 3)
    return K.tocsr(), m


def linear_modal_analysis(X, T, k, bI):
    X = np.asarray(X, float)
    n = X.shape[0]
    K, m = stiffness_and_mass(X, T)
    bI = np.asarray(bI)
    fixed = np.zeros(3 * n, bool)
    for c in range(3):
        fixed[3 * bI + c] = True
    Ii = np.where(~fixed)[0]
    Kf = K[Ii][:, Ii]
    Mf = sps.diags(m[Ii])
    E, Bi = spla.eigsh((Kf + Mf * 1e-6).tocsc(), M=Mf, k=k, sigma=0, which="LM", v0=np.ones(Ii.size))
    order = np.argsort(E)
    B = np.zeros((3 * n, k))
    B[Ii, :] = Bi[:, order]
    return E[order], B
