"""GPU parity: PCG linear solve, Newton / line search / backward Euler drop-ins,
and the device-resident simulator's trajectories against the reference's own
run (10^3 cube, gravity) and the oracle (20^3 cube, prescribed top face).

Position tolerance: 1e-6 relative (max |x - x_ref| / max |x_ref|) after each
kept step, as required for per-step vertex positions.
"""

import os

import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from conftest import GOLDEN, golden, rel_err
from oracle import fem, solver
from oracle.mesh import unit_cube

pytestmark = pytest.mark.gpu


def _quadratic(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    Q = A.T @ A + n * np.eye(n)
    b = rng.standard_normal((n, 1))
    return Q, b, -np.linalg.solve(Q, b)


def test_pcg_solve_matches_direct_solve():
    from simkit_amd.linsolve import pcg_solve
    X, T = unit_cube(5)
    p = solver.BEProblem(X, T, 3e5, 2e6, 1000.0, 1 / 60, [], gravity=-9.8)
    H = p.potential_hessian(X.reshape(-1) + 0.01)
    b = np.random.default_rng(1).standard_normal(H.shape[0])
    x, info = pcg_solve(H, b, rtol=1e-12, return_info=True)
    ref = spla.spsolve(H.tocsc(), b)
    assert rel_err(x, ref) < 1e-9 and info["iters"] > 0
    Q, bq, xs = _quadratic(7, 2)
    assert rel_err(pcg_solve(Q, -bq), xs.ravel()) < 1e-10         # dense input; x* = -Q^-1 b
    with pytest.raises(np.linalg.LinAlgError):
        pcg_solve(np.array([[1.0, 2.0], [2.0, 1.0]]), np.array([1.0, -1.0]))  # p^T A p = -2: breakdown


def test_newton_and_line_search_reference_cases():
    """The reference's own known answers (tests/test_solvers.py:32-54, tests/test_backtracking_line_search.py)."""
    from simkit_amd import backtracking_line_search
    from simkit_amd.solvers import newton_solver
    for n, seed, kw in ((6, 0, dict(max_iter=1, do_line_search=True)), (4, 1, dict(max_iter=5, do_line_search=False))):
        Q, b, xs = _quadratic(n, seed)
        x = newton_solver(np.zeros_like(b), lambda x: float((0.5 * x.T @ Q @ x + b.T @ x).item()),
                          lambda x: Q @ x + b, lambda x: Q, **kw)
        assert np.allclose(x, xs, atol=1e-8)
    Q, b, xs = _quadratic(3, 2)
    x, info = newton_solver(np.zeros_like(b), lambda x: float((0.5 * x.T @ Q @ x + b.T @ x).item()),
                            lambda x: Q @ x + b, lambda x: sps.csr_matrix(Q), max_iter=10, return_info=True)
    assert np.allclose(x, xs, atol=1e-8) and len(info["g"]) == info["iters"] + 1
    q = lambda x: float(0.5 * np.dot(x, x))
    t, x, fx = backtracking_line_search(q, np.array([2.0, -1.0]), np.array([2.0, -1.0]), np.array([-2.0, 1.0]))
    assert t == 1.0 and fx == 0.0
    t, x, fx = backtracking_line_search(q, np.array([0.0]), np.array([1.0]), np.array([1.0]), max_iter=5)
    assert t == 0.0


def test_kinetic_terms_vs_closed_form():
    from simkit_amd.energies import kinetic as kin
    rng = np.random.default_rng(3)
    n = 3000
    x, xc, xp = (rng.standard_normal((n, 1)) for _ in range(3))
    m = rng.uniform(0.5, 2.0, n)
    M = sps.diags(m)
    h = 1 / 60
    xt = xc + h * ((xc - xp) / h)
    d = x - xt
    assert rel_err(kin.be_target(xc, xp, h), xt) < 1e-15
    assert rel_err(kin.velocity_be(xc, xp, h), (xc - xp) / h) < 1e-15
    assert abs(kin.kinetic_energy_be(x, xc, xp, M, h) - 0.5 * (d.ravel() @ (m * d.ravel())) / h ** 2) < 1e-10 * \
        abs(0.5 * (d.ravel() @ (m * d.ravel())) / h ** 2)
    assert rel_err(kin.kinetic_gradient_be(x, xc, xp, M, h), (m[:, None] * d) / h ** 2) < 1e-14
    assert rel_err(kin.kinetic_hessian_be(M, h).diagonal(), m / h ** 2) < 1e-15
    assert kin.kinetic_gradient_be(x, x, x, M, h).shape == (n, 1)


def test_backward_euler_dropin_matches_oracle():
    """simkit_amd.integrators.backward_euler with GPU energies vs the oracle's harness step."""
    import simkit_amd as sk
    from simkit_amd.energies import stable_neo_hookean as snh
    from simkit_amd.integrators import backward_euler
    X, T = unit_cube(4)
    mu, lam = fem.ympr_to_lame(1e6, 0.45)
    J = sk.deformation_jacobian(X, T)
    vol = sk.volume(X, T)
    fixed = np.zeros(X.size, bool)
    for c in range(3):
        fixed[3 * np.where(X[:, 1] == 0)[0] + c] = True
    free = ~fixed
    xfix = X.reshape(-1)[fixed]
    f_ext = sk.gravity_force(X, T, -9.8, 1000.0).reshape(-1)
    m = np.repeat(sk.massmatrix(X, T, 1000.0).diagonal(), 3)

    def expand(z):
        x = np.empty(X.size)
        x[free] = z.ravel()
        x[fixed] = xfix
        return x.reshape(-1, 3)

    E = lambda z: snh.stable_neo_hookean_energy_x(expand(z), J, mu, lam, vol) - f_ext @ expand(z).ravel()
    G = lambda z: (snh.stable_neo_hookean_gradient_x(expand(z), J, mu, lam, vol).ravel() - f_ext)[free][:, None]
    H = lambda z: snh.stable_neo_hookean_hessian_x(expand(z), J, mu, lam, vol)[free][:, free]
    z0 = X.reshape(-1)[free][:, None]
    z1 = backward_euler(z0, z0.copy(), E, G, H, sps.diags(m[free]), 1 / 60, tolerance=1e-10, max_iter=50)
    p = solver.BEProblem(X, T, mu, lam, 1000.0, 1 / 60, np.where(X[:, 1] == 0)[0], gravity=-9.8, newton_rtol=1e-12)
    x1 = p.step(X.reshape(-1), X.reshape(-1), 1 / 60)
    assert np.abs(expand(z1).ravel() - x1).max() <= 1e-6 * np.abs(x1).max()


def _c1_sim(**kw):
    from simkit_amd.sim import BackwardEulerSim
    X, T = unit_cube(10)
    mu, lam = fem.ympr_to_lame(1e6, 0.45)
    return X, BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=1 / 60, material="snh",
                               fixed_vertices=np.where(X[:, 1] == 0)[0], gravity=-9.8, **kw)


def test_sim_cube10_gravity_matches_reference_trajectory(ref_be_cube10):
    X, sim = _c1_sim(pcg_rtol=1e-10)
    kept, infos = sim.rollout(20, keep=set(int(s) for s in ref_be_cube10["steps"]))
    assert all(i["status"] == 0 for i in infos), infos
    for s in ref_be_cube10["steps"]:
        ref = ref_be_cube10[f"x_{s}"]
        err = np.abs(kept[int(s)] - ref).max() / np.abs(ref).max()
        assert err < 1e-6, (s, err)


def test_sim_cube20_prescribed_matches_oracle():
    path = os.path.join(GOLDEN, "oracle_c2_cube20.npz")
    if not os.path.exists(path):
        pytest.skip("oracle_c2_cube20.npz not generated")
    d = golden("oracle_c2_cube20.npz")
    from simkit_amd.sim import BackwardEulerSim, top_face_sinusoid
    X, T = unit_cube(20)
    mu, lam = fem.ympr_to_lame(1e6, 0.45)
    bottom, top = np.where(X[:, 1] == 0.0)[0], np.where(X[:, 1] == 1.0)[0]
    sim = BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=1 / 60, fixed_vertices=np.concatenate([bottom, top]),
                           pcg_rtol=1e-8)
    kept, infos = sim.rollout(20, bc=top_face_sinusoid(X, top, 0.1, 1.0), keep={1, 10, 20})
    assert all(i["status"] == 0 for i in infos), infos
    for s in (1, 10, 20):
        ref = d[f"x_{s}"]
        assert np.abs(kept[s] - ref).max() / np.abs(ref).max() < 1e-6


def test_sim_status_codes_from_host():
    import torch
    from simkit_amd import _lib
    X, sim = _c1_sim(ls_max_iter=0)                     # no trial allowed -> line-search failure
    x = sim.X_dev.clone()
    _, info = sim.step(x, x)
    assert info["status"] == _lib.SK_ERR_LS_FAIL
    X, sim = _c1_sim(newton_max_iter=1)
    _, info = sim.step(x, x)
    assert info["status"] == _lib.SK_ERR_NEWTON_MAXIT and info["newton_iters"] == 1
    bad = x.clone()
    bad[10] = float("nan")
    X, sim = _c1_sim()
    _, info = sim.step(bad, bad)
    assert info["status"] == _lib.SK_ERR_NONFINITE
    assert torch.isfinite(sim.step(x, x)[0]).all()


def test_sim_is_bitwise_deterministic():
    X, sim = _c1_sim()
    a, ia = sim.rollout(3)
    b, ib = sim.rollout(3)
    assert all(np.array_equal(a[k], b[k]) for k in a) and ia == ib


@pytest.mark.parametrize("mat", ["snh", "arap"])
def test_mixed_precision_matches_fp64_oracle(mat):
    """FP32 Hessian / PSD / values / inner PCG with FP64 refinement: per-step positions against the
    FP64 oracle on a jittered 12^3 cube with the prescribed top face (the mixed-precision workload at
    a size the oracle runs in seconds). Newton converges to the same FP64-gradient tolerance, so the
    positions agree to the Newton tolerance level."""
    from simkit_amd.mesh import jitter
    from simkit_amd.sim import BackwardEulerSim, top_face_sinusoid
    n = 12
    X, T = unit_cube(n)
    bottom, top = np.where(X[:, 1] == 0.0)[0], np.where(X[:, 1] == 1.0)[0]
    X = jitter(X, 1.0 / n, 0.2, seed=0, keep=np.concatenate([bottom, top]))
    assert np.all(fem.tet_volumes(X, T) > 0)
    mu, lam = fem.ympr_to_lame(1e6, 0.45)
    P = lambda t: X + np.outer(np.isin(np.arange(X.shape[0]), top), [0.1 * np.sin(2 * np.pi * t), 0, 0])
    steps = 6
    orc = solver.BEProblem(X, T, mu, lam, 1000.0, 1 / 60, np.concatenate([bottom, top]), material=mat,
                           prescribe=P, linsolve="spsolve")
    ref = orc.rollout(steps)
    sim = BackwardEulerSim(X, T, mu, lam, rho=1000.0, h=1 / 60, material=mat,
                           fixed_vertices=np.concatenate([bottom, top]), precision="mixed", pcg_rtol=1e-8)
    kept, infos = sim.rollout(steps, bc=top_face_sinusoid(X, top, 0.1, 1.0))
    assert all(i["status"] == 0 for i in infos), infos
    assert all(i["refine_steps"] >= i["newton_iters"] for i in infos)
    for s in range(1, steps + 1):
        err = np.abs(kept[s] - ref[s - 1]).max() / np.abs(ref[s - 1]).max()
        assert err < 1e-6, (s, err)
