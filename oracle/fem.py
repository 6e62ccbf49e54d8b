"""Oracle FEM element math, PSD projection and assembly (test infrastructure).

FP64 numpy restatement of the reference's per-tet kinematics, stable
neo-Hookean (SNH), ARAP, linear-elasticity, Macklin-Mueller neo-Hookean and
fixed-corotational element tiers, ``psd_project``, and the
``J^T blockdiag(H) J`` Hessian assembly. Layout conventions follow the
reference: ``F`` is row-major (``F[e, i, c]``, i spatial, c material) and
9-vectors / 9x9 blocks index ``vec(F)`` row-major (``3*i + c``).

Independent of the reference's hand-written 81-entry SNH table
(``simkit/energies/stable_neo_hookean.py:333-440``): the SNH Hessian here is
built from the closed form with the cofactor derivative expressed through
Levi-Civita symbols, so a transcription slip in either would show up as a
disagreement with the golden fixture written from the reference.
"""

import numpy as np
import scipy.sparse as sps

# Reference-element shape-function gradient (deformation_jacobian.py:50-54).
_HREF = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])

_EPS = np.zeros((3, 3, 3))
for _i, _j, _k in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
    _EPS[_i, _j, _k] = 1.0
    _EPS[_i, _k, _j] = -1.0


# --------------------------------------------------------------------------- #
# Kinematics and rest measures                                                #
# --------------------------------------------------------------------------- #
def tet_operator(X, T):
    """Per-tet gradient operator D_e = (H Dm^-1)^T, shape (t, 3, 4).

    deformation_jacobian.py:58-61 / deformation_gradient.py:43-62.
    F_e = sum_v x_v D_e[:, v]^T, i.e. F[e, i, c] = sum_v x[T[e,v], i] D[e, c, v].
    """
    XT = X[T].transpose(0, 2, 1)          # (t, 3, 4): columns are the 4 corners
    Dm = XT @ _HREF                       # edge matrix [x1-x0, x2-x0, x3-x0]
    return (_HREF @ np.linalg.inv(Dm)).transpose(0, 2, 1)


def tet_volumes(X, T):
    """Signed volume det([x1-x0, x2-x0, x3-x0]) / 6 (tetrahedron_volumes.py:26-27)."""
    e = X[T[:, 1:]] - X[T[:, [0]]]
    return np.linalg.det(e) / 6.0


def deformation_gradient(D, T, U):
    """F[e] = (D_e U[T_e])^T (deformation_gradient.py:65-67)."""
    return (D @ U[T]).transpose(0, 2, 1)


def deformation_jacobian(D, T, nv):
    """Sparse J (9t x 3n) with J[9e+3i+c, 3v+i] = D_e[c, v] (deformation_jacobian.py:65-86).

    Exact zeros of D are dropped, as the reference's sparse product drops them.
    """
    t = T.shape[0]
    e, i, c, v = np.meshgrid(np.arange(t), np.arange(3), np.arange(3), np.arange(4), indexing="ij")
    rows = 9 * e + 3 * i + c
    cols = 3 * T[e, v] + i
    vals = D[e, c, v]
    J = sps.csc_matrix((vals.ravel(), (rows.ravel(), cols.ravel())), shape=(9 * t, 3 * nv))
    J.eliminate_zeros()
    return J


def lumped_mass(X, T, rho=1.0):
    """Per-vertex lumped mass sum_{e ni v} vol_e*rho/4 (massmatrix.py:40-49)."""
    m = tet_volumes(X, T) * rho
    out = np.zeros(X.shape[0])
    np.add.at(out, T.ravel(), np.repeat(m / T.shape[1], T.shape[1]))
    return out


def gravity_force(X, T, a=-9.8, rho=1.0):
    """M [0, a, 0] per vertex (gravity_force.py:37-41); returns (n, 3)."""
    g = np.zeros_like(X)
    g[:, 1] = lumped_mass(X, T, rho) * a
    return g


def ympr_to_lame(ym, pr):
    """(mu, lam) from Young's modulus and Poisson ratio (ympr_to_lame.py:33-34)."""
    return ym / (2 * (1 + pr)), ym * pr / ((1 + pr) * (1 - 2 * pr))


# --------------------------------------------------------------------------- #
# Stable neo-Hookean element tier                                             #
# --------------------------------------------------------------------------- #
def _cof(F):
    """Cofactor matrix cof_ij = dJ/dF_ij."""
    return 0.5 * np.einsum("imn,jpq,emp,enq->eij", _EPS, _EPS, F, F)


def _dcof(F):
    """d vec(cof) / d vec(F) as (t, 9, 9): d cof_ij / d F_kl = eps_ikn eps_jlq F_nq."""
    d = np.einsum("ikn,jlq,enq->eijkl", _EPS, _EPS, F)
    return d.reshape(-1, 9, 9)


def _snh_scalars(F, mu, lam):
    mu = np.asarray(mu, float).reshape(-1)
    lam = np.asarray(lam, float).reshape(-1)
    IC = np.einsum("eij,eij->e", F, F)
    J = np.linalg.det(F)
    alpha = 1.0 + 3.0 * mu / (4.0 * lam)          # 1 + d mu / ((d+1) lam), d=3 (:88)
    return mu, lam, IC, J, alpha


def snh_energy_element_F(F, mu, lam):
    """psi = mu/2 (I_C-3) - mu/2 log(I_C+1) + lam/2 (J-alpha)^2 (stable_neo_hookean.py:122-126)."""
    mu, lam, IC, J, alpha = _snh_scalars(F, mu, lam)
    return (0.5 * mu * (IC - 3.0) - 0.5 * mu * np.log(IC + 1.0) + 0.5 * lam * (J - alpha) ** 2)[:, None]


def snh_gradient_element_F(F, mu, lam):
    """P = mu I_C/(I_C+1) F + lam (J-alpha) cof(F) (stable_neo_hookean.py:192-215)."""
    mu, lam, IC, J, alpha = _snh_scalars(F, mu, lam)
    A = mu * IC / (IC + 1.0)
    Dc = lam * (J - alpha)
    return A[:, None, None] * F + Dc[:, None, None] * _cof(F)


def snh_hessian_element_F(F, mu, lam):
    """H = A I + B vecF vecF^T + lam vec(cof) vec(cof)^T + lam(J-alpha) dcof/dF (:315-440)."""
    mu, lam, IC, J, alpha = _snh_scalars(F, mu, lam)
    den = IC + 1.0
    A = mu * IC / den
    B = 2.0 * mu / den ** 2
    Dc = lam * (J - alpha)
    f = F.reshape(-1, 9)
    c = _cof(F).reshape(-1, 9)
    H = (A[:, None, None] * np.eye(9)[None]
         + B[:, None, None] * f[:, :, None] * f[:, None, :]
         + lam[:, None, None] * c[:, :, None] * c[:, None, :]
         + Dc[:, None, None] * _dcof(F))
    return H


# --------------------------------------------------------------------------- #
# ARAP element tier                                                           #
# --------------------------------------------------------------------------- #
def svd_rv(F):
    """SVD with the reference's reflection fix (svd_rv.py:33-51, polar_svd.py:34-56).

    L = diag(1, 1, det(U V^T)); U <- U L when det U < 0 < det V, V <- V L when
    det V < 0 < det U; S <- S L. Returns U, sigma (signed, length 3), V.
    """
    U, s, Vt = np.linalg.svd(F)
    V = Vt.transpose(0, 2, 1)
    ldet = np.linalg.det(U @ Vt)
    dU = np.linalg.det(U)
    dV = np.linalg.det(V)
    fixU = (dU < 0) & (dV > 0)
    fixV = (dV < 0) & (dU > 0)
    U = U.copy()
    V = V.copy()
    U[fixU, :, 2] *= ldet[fixU, None]
    V[fixV, :, 2] *= ldet[fixV, None]
    s = s.copy()
    s[:, 2] *= ldet
    return U, s, V


def polar_rotation(F):
    """R = U V^T from svd_rv (polar_svd.py:209-218)."""
    U, _, V = svd_rv(F)
    return U @ V.transpose(0, 2, 1)


# Twist generators and the singular-value pairs they couple (rotation_gradient.py:42-62).
_TWISTS = (
    (np.array([[0, -1, 0], [1, 0, 0], [0, 0, 0]], float), 0, 1),
    (np.array([[0, 0, 0], [0, 0, 1], [0, -1, 0]], float), 1, 2),
    (np.array([[0, 0, 1], [0, 0, 0], [-1, 0, 0]], float), 0, 2),
)


def rotation_gradient_F(F):
    """dR/dF = sum_k 2/max(s_i+s_j, 1e-8) t_k t_k^T, t_k = vec(U T_k V^T)/sqrt2 (rotation_gradient.py:42-72)."""
    U, s, V = svd_rv(F)
    K = np.zeros((F.shape[0], 9, 9))
    for Tk, i, j in _TWISTS:
        t = ((U @ Tk @ V.transpose(0, 2, 1)) / np.sqrt(2.0)).reshape(-1, 9)
        K += (2.0 / np.maximum(s[:, i] + s[:, j], 1e-8))[:, None, None] * t[:, :, None] * t[:, None, :]
    return K


def arap_energy_element_F(F, mu):
    """psi = mu/2 ||F - R||^2 (arap.py:86-91)."""
    mu = np.asarray(mu, float).reshape(-1)
    R = polar_rotation(F)
    return (0.5 * mu * ((F - R) ** 2).sum(axis=(1, 2)))[:, None]


def arap_gradient_element_F(F, mu):
    """P = mu (F - R) (arap.py:109-114)."""
    mu = np.asarray(mu, float).reshape(-1, 1, 1)
    return mu * (F - polar_rotation(F))


def arap_hessian_element_F(F, mu):
    """H = mu (I9 - dR/dF) (arap.py:138-145)."""
    mu = np.asarray(mu, float).reshape(-1, 1, 1)
    return mu * (np.eye(9)[None] - rotation_gradient_F(F))


# --------------------------------------------------------------------------- #
# Linear elasticity, Macklin-Mueller neo-Hookean, fixed corotational          #
# --------------------------------------------------------------------------- #
_TRANSPOSE9 = np.zeros((9, 9))
for _i in range(3):
    for _j in range(3):
        _TRANSPOSE9[3 * _i + _j, 3 * _j + _i] = 1.0
_VECI = np.eye(3).reshape(9)


def linear_energy_element_F(F, mu, lam):
    """psi = mu ||eps||^2 + lam/2 tr(eps)^2, eps = sym(F) - I (linear_elasticity.py, energy element tier)."""
    mu = np.asarray(mu, float).reshape(-1)
    lam = np.asarray(lam, float).reshape(-1)
    eps = 0.5 * (F + F.transpose(0, 2, 1)) - np.eye(3)[None]
    return (mu * (eps ** 2).sum(axis=(1, 2)) + 0.5 * lam * np.trace(eps, axis1=1, axis2=2) ** 2)[:, None]


def linear_gradient_element_F(F, mu, lam):
    """P = mu (F + F^T - 2I) + lam tr(F - I) I (linear_elasticity.py, gradient element tier)."""
    mu = np.asarray(mu, float).reshape(-1, 1, 1)
    lam = np.asarray(lam, float).reshape(-1, 1, 1)
    I = np.eye(3)[None]
    tr = np.trace(F - I, axis1=1, axis2=2)[:, None, None]
    return mu * (F + F.transpose(0, 2, 1) - 2.0 * I) + lam * tr * I


def linear_hessian_element_F(F, mu, lam):
    """H = mu (I + T) + lam vec(I) vec(I)^T, constant in F (linear_elasticity.py, hessian element tier)."""
    mu = np.asarray(mu, float).reshape(-1, 1, 1)
    lam = np.asarray(lam, float).reshape(-1, 1, 1)
    H1 = (np.eye(9) + _TRANSPOSE9)[None]
    H2 = np.outer(_VECI, _VECI)[None]
    return np.broadcast_to(mu * H1 + lam * H2, (F.shape[0], 9, 9)).copy()


def mmnh_energy_element_F(F, mu, lam):
    """psi = mu (1 - J) + lam/2 (1 - J)^2 + mu/2 (I_C - 3) (macklin_mueller_neo_hookean.py, 3D energy)."""
    mu = np.asarray(mu, float).reshape(-1)
    lam = np.asarray(lam, float).reshape(-1)
    c = 1.0 - np.linalg.det(F)
    IC = np.einsum("eij,eij->e", F, F)
    return (mu * c + 0.5 * lam * c ** 2 + 0.5 * mu * (IC - 3.0))[:, None]


def mmnh_gradient_element_F(F, mu, lam):
    """P = mu F - (mu + lam (1 - J)) cof(F)."""
    mu = np.asarray(mu, float).reshape(-1)
    lam = np.asarray(lam, float).reshape(-1)
    d = lam * (np.linalg.det(F) - 1.0) - mu
    return mu[:, None, None] * F + d[:, None, None] * _cof(F)


def mmnh_hessian_element_F(F, mu, lam):
    """H = mu I + lam vec(cof) vec(cof)^T + (lam (J - 1) - mu) dcof/dF."""
    mu = np.asarray(mu, float).reshape(-1)
    lam = np.asarray(lam, float).reshape(-1)
    d = lam * (np.linalg.det(F) - 1.0) - mu
    c = _cof(F).reshape(-1, 9)
    return (mu[:, None, None] * np.eye(9)[None] + lam[:, None, None] * c[:, :, None] * c[:, None, :]
            + d[:, None, None] * _dcof(F))


def fcr_energy_element_F(F, mu, lam):
    """psi = 2 psi_ARAP + lam/2 (J - 1)^2 (fcr.py, energy element tier)."""
    lam = np.asarray(lam, float).reshape(-1)
    return 2.0 * arap_energy_element_F(F, mu) + (0.5 * lam * (np.linalg.det(F) - 1.0) ** 2)[:, None]


def fcr_gradient_element_F(F, mu, lam):
    """P = 2 mu (F - R) + lam (J - 1) cof(F) (fcr.py, gradient element tier)."""
    lam = np.asarray(lam, float).reshape(-1)
    d = lam * (np.linalg.det(F) - 1.0)
    return 2.0 * arap_gradient_element_F(F, mu) + d[:, None, None] * _cof(F)


def fcr_hessian_element_F(F, mu, lam):
    """H = 2 mu (I - dR/dF) + lam vec(cof) vec(cof)^T + lam (J - 1) dcof/dF (fcr.py, hessian element tier)."""
    lam = np.asarray(lam, float).reshape(-1)
    d = lam * (np.linalg.det(F) - 1.0)
    c = _cof(F).reshape(-1, 9)
    Hv = lam[:, None, None] * c[:, :, None] * c[:, None, :] + d[:, None, None] * _dcof(F)
    return 2.0 * arap_hessian_element_F(F, mu) + Hv


# --------------------------------------------------------------------------- #
# PSD projection and assembly                                                 #
# --------------------------------------------------------------------------- #
def psd_project(H, method="proj"):
    """eigh, floor eigenvalues at 1e-6 ('proj') or take |.| ('abs'), rebuild U S U^T (psd_project.py:28-45)."""
    if H.ndim == 2:
        H = H[None]
    s, U = np.linalg.eigh(H)
    if method == "abs":
        s = np.abs(s)
    elif method == "proj":
        s = np.where(s < 1e-6, 1e-6, s)
    return (U * s[:, None, :]) @ U.transpose(0, 2, 1)


def element_stiffness(D, He):
    """12x12 K_e = B_e^T He B_e with B_e[3i+c, 3v+i] = D_e[c, v] (the block of J for tet e)."""
    t = D.shape[0]
    B = np.zeros((t, 9, 12))
    for i in range(3):
        for c in range(3):
            B[:, 3 * i + c, i::3] = D[:, c, :]
    return B.transpose(0, 2, 1) @ He @ B


def assemble_hessian(J, He):
    """Q = J^T blockdiag(He) J, the reference's sparse products (stable_neo_hookean.py:536-537).

    The block diagonal is built as a BSR matrix with one dense 9x9 block per
    element, the same matrix sp.sparse.block_diag(He) builds (explicit zeros
    included); the products then drop exact zeros as the reference's do.
    """
    t = He.shape[0]
    Hb = sps.bsr_matrix((np.ascontiguousarray(He), np.arange(t), np.arange(t + 1)), shape=(9 * t, 9 * t))
    return (J.transpose().tocsr() @ Hb.tocsr() @ J).tocsr()


def assemble_gradient(J, Pe, vol):
    """g = J^T vec(vol P) (stable_neo_hookean.py:500-502)."""
    return J.transpose() @ (Pe * vol.reshape(-1, 1, 1)).reshape(-1, 1)


# --------------------------------------------------------------------------- #
# Global x-tier (vol-weight -> PSD -> assemble; stable_neo_hookean.py:449-538, arap.py:242-325)
# --------------------------------------------------------------------------- #
_ELEMENT = {
    "snh": (snh_energy_element_F, snh_gradient_element_F, snh_hessian_element_F),
    "arap": (lambda F, mu, lam: arap_energy_element_F(F, mu),
             lambda F, mu, lam: arap_gradient_element_F(F, mu),
             lambda F, mu, lam: arap_hessian_element_F(F, mu)),
    "linear": (linear_energy_element_F, linear_gradient_element_F, linear_hessian_element_F),
    "mmnh": (mmnh_energy_element_F, mmnh_gradient_element_F, mmnh_hessian_element_F),
    "fcr": (fcr_energy_element_F, fcr_gradient_element_F, fcr_hessian_element_F),
}


def energy_x(material, x, J, mu, lam, vol):
    F = (J @ x.reshape(-1, 1)).reshape(-1, 3, 3)
    psi = _ELEMENT[material][0](F, mu, lam)
    return float((vol.reshape(-1, 1) * psi).sum())


def element_energies_x(material, x, J, mu, lam, vol):
    """Per-element vol * psi (the summands of energy_x)."""
    F = (J @ x.reshape(-1, 1)).reshape(-1, 3, 3)
    return vol.reshape(-1) * _ELEMENT[material][0](F, mu, lam).reshape(-1)


def gradient_x(material, x, J, mu, lam, vol):
    F = (J @ x.reshape(-1, 1)).reshape(-1, 3, 3)
    return assemble_gradient(J, _ELEMENT[material][1](F, mu, lam), vol)


def projected_element_hessians(material, x, J, mu, lam, vol, psd=True):
    """vol-weighted (then PSD-projected) 9x9 element Hessians (stable_neo_hookean.py:531-535)."""
    F = (J @ x.reshape(-1, 1)).reshape(-1, 3, 3)
    He = _ELEMENT[material][2](F, mu, lam) * vol.reshape(-1, 1, 1)
    return psd_project(He) if psd else He


def hessian_x(material, x, J, mu, lam, vol, psd=True):
    return assemble_hessian(J, projected_element_hessians(material, x, J, mu, lam, vol, psd))
