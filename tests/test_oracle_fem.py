"""Pins the oracle's element math, PSD projection and assembly (CPU only).

Each check is chosen so that a plausible slip (dropped term, wrong sign,
transposed index) fails: golden vectors written from the reference itself,
an independent sympy derivation, finite differences, closed-form invariants
and brute-force assembly.
"""

import numpy as np
import pytest
import scipy.sparse as sps
import sympy as sp

from conftest import per_item_rel_err, rel_err
from oracle import fem
from oracle.mesh import unit_cube


def random_F(seed, n=16, scale=0.3):
    rng = np.random.default_rng(seed)
    return np.eye(3)[None] + scale * rng.standard_normal((n, 3, 3))


# --------------------------------------------------------------------------- #
# Golden vectors from the reference                                            #
# --------------------------------------------------------------------------- #
def test_snh_matches_reference_golden(ref_element):
    d = ref_element
    F, mu, lam = d["F"], d["mu"], d["lam"]
    assert per_item_rel_err(fem.snh_energy_element_F(F, mu, lam), d["snh_psi"]) < 1e-13
    assert per_item_rel_err(fem.snh_gradient_element_F(F, mu, lam), d["snh_P"]) < 1e-13
    assert per_item_rel_err(fem.snh_hessian_element_F(F, mu, lam), d["snh_H"]) < 1e-13


def test_arap_and_svd_branches_match_reference_golden(ref_element):
    d = ref_element
    F, mu = d["F"], d["mu"]
    assert np.sum(np.linalg.det(F) < 0) >= 8, "fixture must exercise the inverted (flip) branch"
    assert per_item_rel_err(fem.arap_energy_element_F(F, mu), d["arap_psi"]) < 1e-12
    assert per_item_rel_err(fem.arap_gradient_element_F(F, mu), d["arap_P"]) < 1e-12
    assert per_item_rel_err(fem.arap_hessian_element_F(F, mu), d["arap_H"]) < 1e-12
    assert per_item_rel_err(fem.rotation_gradient_F(F), d["dRdF"]) < 1e-12


def test_psd_matches_reference_golden(ref_element):
    d = ref_element
    assert per_item_rel_err(fem.psd_project(d["snh_H"]), d["snh_H_psd"]) < 1e-13


def _oracle_assembly(d, name, mat):
    X, T = d["X"], d["T"]
    x = (X if name == "rest" else d["U"]).reshape(-1)
    D = fem.tet_operator(X, T)
    vol = fem.tet_volumes(X, T)
    J = fem.deformation_jacobian(D, T, X.shape[0])
    return fem.hessian_x(mat, x, J, d["mu"], d["lam"], vol), fem.gradient_x(mat, x, J, d["mu"], d["lam"], vol), \
        fem.energy_x(mat, x, J, d["mu"], d["lam"], vol)


@pytest.mark.parametrize("mat", ["snh", "arap"])
def test_assembly_matches_reference_golden(ref_assembly, mat):
    d = ref_assembly
    for name in ("rest", "def"):
        Q, g, E = _oracle_assembly(d, name, mat)
        ref = sps.csr_matrix((d[f"{mat}_{name}_data"], d[f"{mat}_{name}_indices"], d[f"{mat}_{name}_indptr"]),
                             shape=Q.shape)
        assert rel_err(Q.toarray(), ref.toarray()) < 1e-13
        assert rel_err(g, d[f"{mat}_{name}_g"]) < 1e-13
        assert abs(E - float(d[f"{mat}_{name}_E"])) <= 1e-12 * abs(float(d[f"{mat}_{name}_E"])) + 1e-15
        if name == "def":  # generic state: no exact cancellations, pattern is the full block pattern
            Q.sort_indices()
            assert np.array_equal(Q.indptr, ref.indptr) and np.array_equal(Q.indices, ref.indices)


# --------------------------------------------------------------------------- #
# Independent derivation: sympy                                                #
# --------------------------------------------------------------------------- #
@pytest.fixture(scope="module")
def snh_sympy():
    f = sp.symbols("f0:9", real=True)
    mu, lam = sp.symbols("mu lam", positive=True)
    Fm = sp.Matrix(3, 3, f)
    IC = sum(fi ** 2 for fi in f)
    Jd = Fm.det()
    alpha = 1 + sp.Rational(3, 4) * mu / lam
    psi = mu / 2 * (IC - 3) - mu / 2 * sp.log(IC + 1) + lam / 2 * (Jd - alpha) ** 2
    grad = [sp.diff(psi, fi) for fi in f]
    hess = [[sp.diff(gi, fj) for fj in f] for gi in grad]
    return (sp.lambdify((f, mu, lam), psi, "numpy"), sp.lambdify((f, mu, lam), grad, "numpy"),
            sp.lambdify((f, mu, lam), hess, "numpy"))


def test_snh_matches_sympy(snh_sympy):
    e_fn, g_fn, h_fn = snh_sympy
    F = random_F(3, n=6, scale=0.4)
    mu = np.linspace(1e5, 5e5, 6)
    lam = mu * np.linspace(2, 12, 6)
    psi = fem.snh_energy_element_F(F, mu, lam)[:, 0]
    P = fem.snh_gradient_element_F(F, mu, lam)
    H = fem.snh_hessian_element_F(F, mu, lam)
    for k in range(6):
        f = F[k].reshape(-1)
        assert abs(psi[k] - e_fn(f, mu[k], lam[k])) <= 1e-12 * abs(psi[k]) + 1e-9
        assert rel_err(P[k].reshape(-1), np.array(g_fn(f, mu[k], lam[k]), float)) < 1e-12
        assert rel_err(H[k], np.array(h_fn(f, mu[k], lam[k]), float)) < 1e-12


def test_sympy_check_catches_a_dropped_term(snh_sympy, monkeypatch):
    """A planted slip in the oracle (the lam(J-alpha) dcof/dF term removed) must fail the check."""
    _, _, h_fn = snh_sympy
    monkeypatch.setattr(fem, "_dcof", lambda F: np.zeros((F.shape[0], 9, 9)))
    F = random_F(4, n=1)
    H = fem.snh_hessian_element_F(F, 3e5, 2e6)
    assert rel_err(H[0], np.array(h_fn(F[0].reshape(-1), 3e5, 2e6), float)) > 1e-3


# --------------------------------------------------------------------------- #
# Finite differences and invariants                                           #
# --------------------------------------------------------------------------- #
@pytest.mark.parametrize("mat", ["snh", "arap"])
def test_gradient_and_hessian_vs_central_differences(mat):
    F = random_F(5, n=4)
    mu, lam = 3e5, 2e6
    e, g, hfn = fem._ELEMENT[mat]
    P = g(F, mu, lam).reshape(-1, 9)
    H = hfn(F, mu, lam)
    eps = 1e-6
    for a in range(9):
        dF = np.zeros(9)
        dF[a] = eps
        Fp, Fm = F + dF.reshape(3, 3), F - dF.reshape(3, 3)
        fd_P = (e(Fp, mu, lam) - e(Fm, mu, lam))[:, 0] / (2 * eps)
        assert np.allclose(fd_P, P[:, a], rtol=1e-6, atol=1e-3)
        fd_H = (g(Fp, mu, lam) - g(Fm, mu, lam)).reshape(-1, 9) / (2 * eps)
        assert np.allclose(fd_H, H[:, a, :], rtol=1e-5, atol=1e-1)


def test_snh_rest_state():
    F = np.eye(3)[None]
    assert np.abs(fem.snh_gradient_element_F(F, 3e5, 2e6)).max() < 1e-9
    H = fem.snh_hessian_element_F(F, 3e5, 2e6)[0]
    assert np.linalg.eigvalsh(H).min() > -1e-6 * np.abs(H).max()
    assert np.allclose(H, H.T)


def test_svd_rv_invariants():
    F = random_F(6, n=50, scale=0.8)
    U, s, V = fem.svd_rv(F)
    R = fem.polar_rotation(F)
    assert np.allclose(np.linalg.det(R), 1.0)
    assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3)[None], atol=1e-12)
    S = V @ (s[:, :, None] * np.eye(3)[None]) @ V.transpose(0, 2, 1)
    assert np.allclose(R @ S, F, atol=1e-12)
    assert np.all(np.sign(s[:, 2]) == np.sign(np.linalg.det(F)))


def test_arap_twist_eigenstructure():
    """H = mu (I - dR/dF): eigenvalue mu on the 6 non-twist directions, mu (1 - 2/(s_i+s_j)) on twists."""
    F = random_F(7, n=8, scale=0.2)
    mu = 2.5e5
    H = fem.arap_hessian_element_F(F, mu)
    _, s, _ = fem.svd_rv(F)
    for k in range(8):
        ev = np.sort(np.linalg.eigvalsh(H[k]))
        tw = np.sort([mu * (1 - 2 / (s[k, 0] + s[k, 1])), mu * (1 - 2 / (s[k, 1] + s[k, 2])),
                      mu * (1 - 2 / (s[k, 0] + s[k, 2]))])
        assert np.allclose(ev, np.sort(np.concatenate([tw, np.full(6, mu)])), rtol=1e-10, atol=1e-6)


def test_rotation_gradient_clamp_spectrum():
    """Where s_i + s_j < 1e-8 (rotation_gradient.py:65-67 clamps it to 1e-8, also when negative for
    inverted F) dR/dF keeps the twist eigenstructure: eigenvalues 2/max(s_i + s_j, 1e-8) on the three
    twists, 0 on the other six directions (eigvalsh, independent of the outer-product form)."""
    rng = np.random.default_rng(8)
    for sig in ((1.0, 3e-9, 2e-9), (1.0, 3e-9, -2e-9), (1.0, 1e-9, -2e-9), (1.0, 1e-9, 0.0), (0.7, 1.2e-8, 5e-9)):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        q *= np.sign(np.linalg.det(q))
        F = (q @ np.diag(sig))[None]
        K = fem.rotation_gradient_F(F)[0]
        s = np.sort(np.abs(sig))[::-1] * np.array([1, 1, np.sign(np.prod(sig)) or 1.0])
        want = np.sort(np.concatenate([np.zeros(6), [2 / max(s[0] + s[1], 1e-8), 2 / max(s[1] + s[2], 1e-8),
                                                     2 / max(s[0] + s[2], 1e-8)]]))
        ev = np.sort(np.linalg.eigvalsh(K))
        assert np.allclose(ev, want, rtol=1e-9, atol=1e-6 * want.max()), (sig, ev, want)


def test_psd_invariants():
    rng = np.random.default_rng(8)
    A = rng.standard_normal((20, 9, 9))
    A = A + A.transpose(0, 2, 1)
    P = fem.psd_project(A)
    assert np.linalg.eigvalsh(P).min() >= 1e-6 - 1e-12
    # commutes with the input and leaves already-PD matrices unchanged
    assert np.allclose(P @ A, A @ P, atol=1e-10)
    B = A @ A.transpose(0, 2, 1) + np.eye(9)
    assert np.allclose(fem.psd_project(B), B, atol=1e-12)
    Pa = fem.psd_project(A, "abs")
    assert np.allclose(np.sort(np.linalg.eigvalsh(Pa)), np.sort(np.abs(np.linalg.eigvalsh(A))), atol=1e-10)


# --------------------------------------------------------------------------- #
# Kinematics and assembly                                                     #
# --------------------------------------------------------------------------- #
def test_affine_maps_recovered_exactly():
    X, T = unit_cube(3)
    D = fem.tet_operator(X, T)
    rng = np.random.default_rng(9)
    for _ in range(3):
        A = rng.standard_normal((3, 3))
        b = rng.standard_normal(3)
        F = fem.deformation_gradient(D, T, X @ A.T + b)
        assert np.abs(F - A[None]).max() < 1e-12
    J = fem.deformation_jacobian(D, T, X.shape[0])
    U = X + 0.1 * rng.standard_normal(X.shape)
    assert np.abs((J @ U.reshape(-1)).reshape(-1, 3, 3) - fem.deformation_gradient(D, T, U)).max() < 1e-12


def test_assembly_equals_dense_brute_force():
    X, T = unit_cube(2)
    rng = np.random.default_rng(10)
    x = (X + 0.05 * rng.standard_normal(X.shape)).reshape(-1)
    D = fem.tet_operator(X, T)
    vol = fem.tet_volumes(X, T)
    J = fem.deformation_jacobian(D, T, X.shape[0])
    He = fem.projected_element_hessians("snh", x, J, 3e5, 2e6, vol)
    Ke = fem.element_stiffness(D, He)
    n = 3 * X.shape[0]
    Q = np.zeros((n, n))
    for e in range(T.shape[0]):
        dofs = (3 * T[e][:, None] + np.arange(3)[None]).reshape(-1)
        Q[np.ix_(dofs, dofs)] += Ke[e]
    assert rel_err(fem.hessian_x("snh", x, J, 3e5, 2e6, vol).toarray(), Q) < 1e-13


def test_lumped_mass_and_gravity():
    X, T = unit_cube(4)
    m = fem.lumped_mass(X, T, rho=1000.0)
    assert np.isclose(m.sum(), 1000.0)
    g = fem.gravity_force(X, T, -9.8, 1000.0)
    assert np.isclose(g[:, 1].sum(), -9800.0) and np.all(g[:, [0, 2]] == 0)
