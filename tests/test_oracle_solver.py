"""Pins the oracle's line search, Newton and backward-Euler harness (CPU only)."""

import numpy as np
import pytest
import scipy.sparse as sps

from oracle import fem, solver
from oracle.mesh import unit_cube


# The reference's own known-answer tests (tests/test_backtracking_line_search.py:15-54).
def test_line_search_full_step_on_quadratic():
    q = lambda x: float(0.5 * np.dot(x, x))
    x0 = np.array([2.0, -1.0])
    t, x, fx = solver.backtracking_line_search(q, x0, x0.copy(), -x0)
    assert t == 1.0 and np.allclose(x, 0) and fx == 0.0


def test_line_search_armijo_holds():
    f = lambda x: float(x[0] ** 4)
    x0, g, dx = np.array([1.5]), np.array([3.0]), np.array([-1.0])
    t, x, fx = solver.backtracking_line_search(f, x0, g, dx, alpha=0.01, beta=0.5, max_iter=50)
    assert fx <= f(x0) + 0.01 * t * (g @ dx) + 1e-12 and np.allclose(x, x0 + t * dx)


def test_line_search_zero_step_on_ascent():
    q = lambda x: float(0.5 * np.dot(x, x))
    t, x, fx = solver.backtracking_line_search(q, np.array([0.0]), np.array([1.0]), np.array([1.0]), max_iter=5)
    assert t == 0.0 and np.allclose(x, 0.0)


# tests/test_solvers.py:32-54: Newton is exact in one step on an SPD quadratic.
def test_newton_exact_on_quadratic():
    rng = np.random.default_rng(0)
    A = rng.standard_normal((6, 6))
    Q = A.T @ A + 6 * np.eye(6)
    b = rng.standard_normal((6, 1))
    x_star = -np.linalg.solve(Q, b)
    x = solver.newton_solver(np.zeros_like(b), lambda x: float((0.5 * x.T @ Q @ x + b.T @ x).item()),
                             lambda x: Q @ x + b, lambda x: sps.csr_matrix(Q), max_iter=1)
    assert np.allclose(x, x_star, atol=1e-8)


def test_block_jacobi_pcg_matches_direct_solve():
    X, T = unit_cube(3)
    p = solver.BEProblem(X, T, 3e5, 2e6, 1000.0, 1 / 60, np.where(X[:, 1] == 0)[0], gravity=-9.8)
    x = X.reshape(-1) + 0.01 * np.random.default_rng(1).standard_normal(X.size)
    H = p.potential_hessian(x)
    b = np.random.default_rng(2).standard_normal(X.size)
    fr = p.free
    dx, it = solver.block_jacobi_pcg(H, b, fr, rtol=1e-12)
    ref = np.zeros_like(b)
    import scipy.sparse.linalg as spla
    ref[fr] = spla.spsolve(H[fr][:, fr].tocsc(), b[fr])
    assert np.abs(dx - ref).max() <= 1e-9 * np.abs(ref).max() and it > 0


def test_be_harmonic_oscillator_closed_form():
    """One-DOF spring with BE: x_{n+1} = (x_n + h v_n) / (1 + h^2 k/m), the textbook recurrence."""
    k, m, h = 40.0, 2.0, 0.05
    # minimise 0.5 k x^2 + m/(2h^2)(x - x~)^2 with the oracle's Newton on a 1x1 system
    xc, xp = 1.0, 1.0
    for _ in range(10):
        xt = solver.be_target(np.array([xc]), np.array([xp]), h)
        x = solver.newton_solver(xt.reshape(-1, 1).copy(),
                                 lambda x: float(0.5 * k * x[0, 0] ** 2 + m / (2 * h * h) * (x[0, 0] - xt[0]) ** 2),
                                 lambda x: np.array([[k * x[0, 0] + m / (h * h) * (x[0, 0] - xt[0])]]),
                                 lambda x: np.array([[k + m / (h * h)]]), max_iter=5)[0, 0]
        v = (xc - xp) / h
        assert np.isclose(x, (xc + h * v) / (1 + h * h * k / m), rtol=1e-12)
        xp, xc = xc, x


def test_be_cube_matches_reference_trajectory(ref_be_cube10):
    """Oracle harness vs the reference's own backward_euler run (first steps; 1e-6 relative bar)."""
    X, T = unit_cube(10)
    mu, lam = fem.ympr_to_lame(1e6, 0.45)
    p = solver.BEProblem(X, T, mu, lam, 1000.0, 1 / 60, np.where(X[:, 1] == 0)[0], gravity=-9.8)
    states, infos = p.rollout(2, return_info=True)
    assert all(i["status"] == "ok" for i in infos)
    for s in (1, 2):
        ref = ref_be_cube10[f"x_{s}"]
        assert np.abs(states[s - 1] - ref).max() <= 1e-6 * np.abs(ref).max()


def test_harness_stops_when_updates_fall_below_position_resolution():
    """A stiff short cantilever damps to rest under gravity; from then on ||g(x~)|| is tiny and the
    relative test would demand sub-roundoff gradients. The harness must stop once an accepted
    update no longer moves x (4 eps max |x|) instead of running to the iteration cap."""
    from simkit_amd.mesh import beam
    X, T = beam(8, 4, 0.025)
    mu, lam = fem.ympr_to_lame(6.22e6, 0.479)
    p = solver.BEProblem(X, T, mu, lam, 1000.0, 1 / 60, np.where(X[:, 0] == 0)[0], gravity=-9.8,
                         newton_max_iter=40)
    _, infos = p.rollout(12, return_info=True)
    assert all(i["status"] == "ok" for i in infos)
    assert any(i["stalled"] for i in infos) and max(i["newton"] for i in infos) <= 6
