"""Pins the oracle's linear-elasticity, Macklin-Mueller neo-Hookean and
fixed-corotational element tiers and its BDF2 harness (CPU only).

Golden vectors written from the reference's own element tables
(tests/golden/ref_materials.npz), central differences of energy -> P -> H,
rest-state invariants, and the BDF2 step's exactness on a parabola.
"""

import numpy as np
import pytest

from conftest import golden, per_item_rel_err
from oracle import fem, solver
from oracle.mesh import unit_cube

MATS = ["linear", "mmnh", "fcr"]


def random_F(seed, n=16, scale=0.3):
    rng = np.random.default_rng(seed)
    return np.eye(3)[None] + scale * rng.standard_normal((n, 3, 3))


@pytest.fixture(scope="module")
def ref_materials():
    return golden("ref_materials.npz")


@pytest.mark.parametrize("mat", MATS)
def test_matches_reference_golden(ref_materials, mat):
    d = ref_materials
    e, g, h = fem._ELEMENT[mat]
    assert per_item_rel_err(e(d["F"], d["mu"], d["lam"]), d[f"{mat}_psi"]) < 1e-13
    assert per_item_rel_err(g(d["F"], d["mu"], d["lam"]), d[f"{mat}_P"]) < 1e-13
    assert per_item_rel_err(h(d["F"], d["mu"], d["lam"]), d[f"{mat}_H"]) < 1e-13


def test_golden_catches_a_dropped_term(ref_materials):
    """A plausible slip in the oracle (the MM-NH cofactor coefficient losing its -mu) fails the golden check."""
    d = ref_materials
    c = fem._cof(d["F"])
    P_bad = fem.mmnh_gradient_element_F(d["F"], d["mu"], d["lam"]) + d["mu"].reshape(-1, 1, 1) * c
    assert per_item_rel_err(P_bad, d["mmnh_P"]) > 1e-3


@pytest.mark.parametrize("mat", MATS)
def test_gradient_and_hessian_vs_central_differences(mat):
    F = random_F(25, n=6, scale=0.4)
    F[0, :, 0] *= -1.0                     # one inverted element
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
    assert np.allclose(H, H.transpose(0, 2, 1))


@pytest.mark.parametrize("mat", MATS)
def test_rest_state_is_stress_free(mat):
    e, g, h = fem._ELEMENT[mat]
    F = np.eye(3)[None]
    assert abs(e(F, 3e5, 2e6)[0, 0]) < 1e-9
    assert np.abs(g(F, 3e5, 2e6)).max() < 1e-9


def test_linear_hessian_closed_form():
    H = fem.linear_hessian_element_F(random_F(1, n=3), 2.0, 5.0)
    assert np.allclose(H[0], H[2])                      # constant in F
    w = np.linalg.eigvalsh(H[0])
    # sym(dF) directions: 2 mu (x5 deviatoric) and 2 mu + 3 lam (dilation); skew: 0 (x3)
    assert np.allclose(np.sort(w), np.sort([0, 0, 0, 4, 4, 4, 4, 4, 4 + 15]))


def test_fcr_is_twice_arap_plus_volume_term():
    F = random_F(3, n=20, scale=0.5)
    mu, lam = 1.7, 3.1
    J = np.linalg.det(F)
    assert np.allclose(fem.fcr_energy_element_F(F, mu, lam)[:, 0],
                       2 * fem.arap_energy_element_F(F, mu)[:, 0] + 0.5 * lam * (J - 1) ** 2)


def test_bdf2_target_reproduces_a_parabola():
    """For x(t) = a + b t + c t^2 sampled at t_k, the BDF2 step x~ + (4h^2/9) x'' is exact (BDF2 is
    second order), which fixes both the target and the 9/4 kinetic coefficient."""
    rng = np.random.default_rng(4)
    a, b, c = rng.standard_normal((3, 12))
    h = 0.05
    x = lambda t: a + b * t + c * t ** 2
    t0 = 0.3
    xt = solver.bdf2_target(x(t0), x(t0 - h), x(t0 - 2 * h), x(t0 - 3 * h), h)
    assert np.allclose(xt + (4 * h * h / 9.0) * 2 * c, x(t0 + h), rtol=1e-13, atol=1e-13)


def test_bdf2_harness_free_fall_recurrence():
    """Unpinned mesh under gravity: elastic forces vanish on rigid translation, so every vertex follows
    x_{n+1} = x~_n + (4h^2/9) g exactly (the harness's c and target in the Newton problem)."""
    X, T = unit_cube(2)
    mu, lam = fem.ympr_to_lame(1e5, 0.3)
    h = 1 / 60
    p = solver.BEProblem(X, T, mu, lam, 1000.0, h, [], gravity=-9.8, integrator="bdf2", newton_rtol=1e-12)
    states = p.rollout(5)
    y = [0.0] * 4
    for n in range(5):
        yt = solver.bdf2_target(y[0], y[1], y[2], y[3], h)
        y = [yt + 4 * h * h / 9.0 * -9.8] + y[:3]
        assert np.allclose(states[n].reshape(-1, 3)[:, 1] - X[:, 1], y[0], rtol=1e-9, atol=1e-12)
        assert np.allclose(states[n].reshape(-1, 3)[:, [0, 2]], X[:, [0, 2]], atol=1e-12)
