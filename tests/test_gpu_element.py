"""GPU parity: element tier, PSD projection, SVD / rotation gradient and rest
measures against the oracle and the reference's golden vectors.

Tolerance: 1e-10 relative, measured per element on that element's own scale
(max |gpu - ref| / max |ref| over each element's block), in FP64.
"""

import numpy as np
import pytest

from conftest import per_item_rel_err, rel_err
from oracle import fem
from oracle.mesh import unit_cube

pytestmark = pytest.mark.gpu
TOL = 1e-10


def random_F(seed, n=256, scale=0.3, inverted=0):
    rng = np.random.default_rng(seed)
    F = np.eye(3)[None] + scale * rng.standard_normal((n, 3, 3))
    F[:inverted, :, 0] *= -1.0
    return F


def test_snh_element_tier_vs_golden_and_oracle(ref_element):
    from simkit_amd.energies import stable_neo_hookean as snh
    d = ref_element
    F, mu, lam = d["F"], d["mu"], d["lam"]
    assert per_item_rel_err(snh.stable_neo_hookean_energy_element_F(F, mu, lam), d["snh_psi"]) < TOL
    assert per_item_rel_err(snh.stable_neo_hookean_gradient_element_F(F, mu, lam), d["snh_P"]) < TOL
    assert per_item_rel_err(snh.stable_neo_hookean_hessian_element_F(F, mu, lam), d["snh_H"]) < TOL
    F2 = random_F(11, 1000, 0.5, inverted=100)
    assert per_item_rel_err(snh.stable_neo_hookean_hessian_element_F(F2, 3e5, 2e6),
                            fem.snh_hessian_element_F(F2, 3e5, 2e6)) < TOL
    assert per_item_rel_err(snh.stable_neo_hookean_energy_element_F(F2, 3e5, 2e6),
                            fem.snh_energy_element_F(F2, 3e5, 2e6)) < TOL


def test_arap_element_tier_vs_golden_and_oracle(ref_element):
    from simkit_amd.energies import arap
    d = ref_element
    F, mu = d["F"], d["mu"]
    assert per_item_rel_err(arap.arap_energy_element_F(F, mu), d["arap_psi"]) < TOL
    assert per_item_rel_err(arap.arap_gradient_element_F(F, mu), d["arap_P"]) < TOL
    assert per_item_rel_err(arap.arap_hessian_element_F(F, mu), d["arap_H"]) < TOL
    F2 = random_F(12, 1000, 0.4, inverted=100)
    assert per_item_rel_err(arap.arap_hessian_element_F(F2, 2e5), fem.arap_hessian_element_F(F2, 2e5)) < TOL
    assert per_item_rel_err(arap.arap_gradient_element_F(F2, 2e5), fem.arap_gradient_element_F(F2, 2e5)) < TOL


def test_rotation_gradient_and_polar(ref_element):
    import simkit_amd as sk
    d = ref_element
    assert per_item_rel_err(sk.rotation_gradient_F(d["F"]), d["dRdF"]) < TOL
    F = random_F(13, 500, 0.8, inverted=60)
    R, S = sk.polar_svd(F)
    assert rel_err(R, fem.polar_rotation(F)) < TOL
    assert np.allclose(np.linalg.det(R), 1.0, atol=1e-12)
    assert rel_err(R @ S, F) < 1e-12
    U, Sd, V = sk.svd_rv(F)
    assert rel_err(U @ Sd @ V.transpose(0, 2, 1), F) < 1e-12
    assert np.all(np.linalg.det(U @ V.transpose(0, 2, 1)) > 0)


def test_psd_project_vs_oracle(ref_element):
    import simkit_amd as sk
    d = ref_element
    assert per_item_rel_err(sk.psd_project(d["snh_H"]), d["snh_H_psd"]) < TOL
    rng = np.random.default_rng(14)
    for dim in (1, 2, 3, 7, 8, 9, 10, 11, 12, 13, 20):   # > 12: the scratch-memory Jacobi
        A = rng.standard_normal((300, dim, dim))
        A = A + A.transpose(0, 2, 1)
        for method in ("proj", "abs"):
            assert per_item_rel_err(sk.psd_project(A, method), fem.psd_project(A, method)) < TOL
    # the reference's own cases (tests/test_psd_project.py:10-23)
    P = sk.psd_project(np.array([[1.0, 0.5], [0.5, -0.2]]))[0]
    assert np.all(np.linalg.eigvalsh(P) >= 1e-6 - 1e-12)
    # vol-weighted SNH Hessians on a deformed mesh: the 9x9 projection the x tier uses
    F = random_F(15, 2000, 0.35, inverted=50)
    H = fem.snh_hessian_element_F(F, 3.4e5, 3.1e6) * 1.2e-6
    assert per_item_rel_err(sk.psd_project(H), fem.psd_project(H)) < TOL
    # lower triangle only (as numpy.linalg.eigh), in registers (12) and in scratch memory (20)
    for dim in (12, 20):
        A = rng.standard_normal((100, dim, dim))
        A = A + A.transpose(0, 2, 1)
        B = A.copy()
        iu = np.triu_indices(dim, 1)
        B[:, iu[0], iu[1]] = 7.0
        assert per_item_rel_err(sk.psd_project(B), fem.psd_project(A)) < TOL


def test_rest_measures_and_kinematics():
    import simkit_amd as sk
    X, T = unit_cube(6)
    rng = np.random.default_rng(16)
    Xj = X + 0.02 * rng.standard_normal(X.shape)
    assert rel_err(sk.tetrahedron_volumes(Xj, T), fem.tet_volumes(Xj, T)) < 1e-12
    assert sk.volume(Xj, T).shape == (T.shape[0], 1)
    M = sk.massmatrix(Xj, T, rho=1000.0)
    assert rel_err(M.diagonal(), fem.lumped_mass(Xj, T, 1000.0)) < 1e-12
    assert rel_err(sk.gravity_force(Xj, T, -9.8, 1000.0), fem.gravity_force(Xj, T, -9.8, 1000.0)) < 1e-12
    U = Xj + 0.05 * rng.standard_normal(X.shape)
    D = fem.tet_operator(Xj, T)
    assert rel_err(sk.deformation_gradient(Xj, T, U), fem.deformation_gradient(D, T, U)) < 1e-12
    J = sk.deformation_jacobian(Xj, T)
    Jo = fem.deformation_jacobian(D, T, X.shape[0])
    assert J.shape == Jo.shape and J.format == "csc"
    assert rel_err(J.toarray(), Jo.toarray()) < 1e-12
    A = rng.standard_normal((3, 3))
    F = (J @ (Xj @ A.T).reshape(-1)).reshape(-1, 3, 3)
    assert np.abs(F - A[None]).max() < 1e-10
    mu, lam = sk.ympr_to_lame(1e6, 0.45)
    assert np.isclose(mu, 1e6 / 2.9) and np.isclose(lam, 1e6 * 0.45 / (1.45 * 0.1))
