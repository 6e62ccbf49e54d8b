"""Oracle line search, Newton and backward-Euler harness (test infrastructure).

``backtracking_line_search`` and ``newton_solver`` restate the reference
(``simkit/backtracking_line_search.py:8-67``, ``simkit/solvers/newton.py:7-75``)
including its step-norm stop. ``BEProblem`` is the benchmark harness: the
reference's backward-Euler incremental potential
(``simkit/integrators/backward_euler.py:76-91``, ``simkit/energies/kinetic.py:87-194``)
minimised by Newton with a PSD-projected Hessian and Armijo backtracking, with
two explicit harness choices that the reference leaves to its callers:

* Pinned / prescribed vertices are eliminated (their DOFs are held at the
  prescribed values, as ``simkit/linear_modal_analysis.py:52-66`` eliminates
  DOFs) rather than penalised.
* Newton stops on ``||g_free(x_k)|| <= rtol * ||g_free(x_0)||`` (the reference's
  own loop stops on ``||alpha dx||``; its timing demo unrolls the loop the same
  way, ``examples/interactive_demos/003_resolution_complexity.py:113-155``).
* The Armijo test ``f(x0+t dx) <= f(x0) + a t g.dx + 1e-12`` is evaluated in
  difference form ``sum_e vol_e (psi_e(x0+t dx) - psi_e(x0)) + (kinetic and
  load differences) <= a t g.dx + 1e-12``. Algebraically identical; it avoids
  cancelling two ~1e5 sums whose rounding (~1e-11) exceeds the 1e-12 slack.
"""

import time

import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla

from . import fem


def backtracking_line_search(f, x0, g, dx, alpha=0.01, beta=0.5, max_iter=100, threshold=1e-12):
    """backtracking_line_search.py:52-67."""
    assert 0 < alpha <= 0.5
    assert 0 < beta < 1
    assert np.ndim(x0) == np.ndim(dx)
    t = 1.0
    fx0 = f(x0)
    for _ in range(max_iter):
        x = x0 + t * dx
        fx = f(x)
        if fx <= fx0 + alpha * t * (g.T @ dx) + threshold:
            return t, x, fx
        t = beta * t
    return 0.0, x0, fx0


def newton_solver(x0, energy_func, gradient_func, hessian_func, tolerance=1e-6, max_iter=1,
                  do_line_search=True, return_info=False):
    """newton.py:42-75 (direct solve; step-norm stop)."""
    x = x0.copy()
    info = {"g": [], "dx": [], "alphas": [], "iters": -1}
    for i in range(max_iter):
        g = gradient_func(x)
        H = hessian_func(x)
        if sps.issparse(H):
            dx = spla.spsolve(H.tocsc(), -g).reshape(-1, 1)
        else:
            dx = np.linalg.solve(H, -g).reshape(-1, 1)
        alpha = backtracking_line_search(energy_func, x, g, dx)[0] if do_line_search else 1.0
        x += alpha * dx
        info["g"].append(g)
        info["dx"].append(dx)
        info["alphas"].append(alpha)
        info["iters"] = i
        if np.linalg.norm(alpha * dx) < tolerance:
            break
    return (x, info) if return_info else x


def be_target(x_curr, x_prev, h):
    """x_curr + h (x_curr - x_prev)/h, computed as the reference does (kinetic.py:87-90)."""
    return x_curr + h * ((x_curr - x_prev) / h)


def bdf2_target(x_curr, x_prev, x_prev2, x_prev3, h):
    """BDF2 target (kinetic.py bdf2_target / velocity_bdf2), same order of operations:
    (4/3) x_c - (1/3) x_p + (8h/9) v_c - (2h/9) v_p, v = (3 x_k - 4 x_k-1 + x_k-2) / (2h)."""
    v_curr = (3.0 * x_curr - 4.0 * x_prev + x_prev2) / (2.0 * h)
    v_prev = (3.0 * x_prev - 4.0 * x_prev2 + x_prev3) / (2.0 * h)
    return (4.0 / 3.0) * x_curr - (1.0 / 3.0) * x_prev + (8.0 * h / 9.0) * v_curr - (2.0 * h / 9.0) * v_prev


def block_jacobi_pcg(A, b, free, rtol=1e-8, maxit=10000):
    """CPU block-Jacobi PCG on the free DOFs (same algorithm the GPU runs; for baselines)."""
    n = A.shape[0]
    nb = n // 3
    diag = np.zeros((nb, 3, 3))
    Acoo = A.tocoo()
    m = (Acoo.row // 3) == (Acoo.col // 3)
    diag[Acoo.row[m] // 3, Acoo.row[m] % 3, Acoo.col[m] % 3] = Acoo.data[m]
    fm = free.reshape(nb, 3)
    diag = diag * fm[:, :, None] * fm[:, None, :] + np.eye(3)[None] * (~fm)[:, :, None]
    Binv = np.linalg.inv(diag)
    fr = free.astype(float)

    def prec(r):
        return (Binv @ r.reshape(nb, 3, 1)).reshape(-1) * fr

    x = np.zeros(n)
    r = b * fr
    z = prec(r)
    p = z.copy()
    rz = r @ z
    r0 = np.sqrt(r @ r)
    it = 0
    while np.sqrt(r @ r) > rtol * r0 and it < maxit:
        Ap = (A @ p) * fr
        a = rz / (p @ Ap)
        x += a * p
        r -= a * Ap
        z = prec(r)
        rz_new = r @ z
        p = z + (rz_new / rz) * p
        rz = rz_new
        it += 1
    return x, it


class BEProblem:
    """Backward-Euler harness on a tet mesh with eliminated (pinned/prescribed) DOFs.

    E(x) = V(x) + c/(2h^2) (x - x~)^T M (x - x~) - f_ext^T x, minimised over the
    free DOFs of x (backward_euler.py:76-91 with the kinetic terms of
    kinetic.py:107-194 and the gravity load of gravity_force.py:37-41).
    integrator="bdf2" uses c = 9/4 and the BDF2 target (bdf2.py:86-101).
    Newton stops at ||g_free|| <= newton_rtol ||g_free(x0)||, or when an accepted
    update is below the position resolution (max |t dx| <= 4 eps max |x|): x0
    was then already an equilibrium to within roundoff and the relative test
    cannot be met (newton.py:69 likewise stops on the step length).
    """

    def __init__(self, X, T, mu, lam, rho, h, fixed_vertices, material="snh",
                 gravity=0.0, prescribe=None, linsolve="spsolve", newton_rtol=1e-6,
                 newton_max_iter=100, pcg_rtol=1e-8, integrator="be"):
        self.X = np.asarray(X, float)
        self.T = np.asarray(T)
        self.nv = X.shape[0]
        self.material = material
        t = self.T.shape[0]
        self.mu = np.broadcast_to(np.asarray(mu, float), (t,)).copy()
        self.lam = np.broadcast_to(np.asarray(lam, float), (t,)).copy()
        self.h = h
        self.D = fem.tet_operator(self.X, self.T)
        self.vol = fem.tet_volumes(self.X, self.T)
        self.J = fem.deformation_jacobian(self.D, self.T, self.nv)
        self.m = np.repeat(fem.lumped_mass(self.X, self.T, rho), 3)
        self.f_ext = fem.gravity_force(self.X, self.T, gravity, rho).reshape(-1) if gravity else np.zeros(3 * self.nv)
        self.fixed = np.zeros(3 * self.nv, bool)
        fv = np.asarray(fixed_vertices, dtype=np.int64).reshape(-1)
        for c in range(3):
            self.fixed[3 * fv + c] = True
        self.free = ~self.fixed
        self.prescribe = prescribe          # callable(time) -> (nv,3) targets for fixed vertices, or None
        self.linsolve = linsolve
        self.newton_rtol = newton_rtol
        self.newton_max_iter = newton_max_iter
        self.pcg_rtol = pcg_rtol
        self.integrator = integrator
        self.c = {"be": 1.0, "bdf2": 9.0 / 4.0}[integrator]
        # optional per-phase wall times (seconds), accumulated by step() when set to a dict
        self.timers = None

    def _tick(self, key, t0):
        t1 = time.perf_counter()
        if self.timers is not None:
            self.timers[key] = self.timers.get(key, 0.0) + (t1 - t0)
        return t1

    # -- incremental potential pieces -------------------------------------
    def potential_gradient(self, x, xt):
        g = fem.gradient_x(self.material, x, self.J, self.mu, self.lam, self.vol).reshape(-1)
        return g + self.c * self.m * (x - xt) / self.h ** 2 - self.f_ext

    def potential_hessian(self, x):
        Q = fem.hessian_x(self.material, x, self.J, self.mu, self.lam, self.vol, psd=True)
        return (Q + sps.diags(self.c * self.m / self.h ** 2)).tocsr()

    def energy_delta(self, x0, x1, xt):
        """E(x1) - E(x0) summed as per-element differences (see module docstring)."""
        e1 = fem.element_energies_x(self.material, x1, self.J, self.mu, self.lam, self.vol)
        e0 = fem.element_energies_x(self.material, x0, self.J, self.mu, self.lam, self.vol)
        d0 = x0 - xt
        d1 = x1 - xt
        kin = 0.5 * self.c / self.h ** 2 * (self.m * (d1 - d0) * (d1 + d0)).sum()
        return (e1 - e0).sum() + kin - self.f_ext @ (x1 - x0)

    def step(self, x_curr, x_prev, t_next, return_info=False, x_prev2=None, x_prev3=None):
        """One BE (or BDF2) step; returns x_next (and a dict of per-step counters)."""
        if self.integrator == "bdf2":
            xt = bdf2_target(x_curr, x_prev, x_prev2, x_prev3, self.h)
        else:
            xt = be_target(x_curr, x_prev, self.h)
        x = xt.copy()
        if self.prescribe is not None:
            x[self.fixed] = self.prescribe(t_next).reshape(-1)[self.fixed]
        else:
            x[self.fixed] = x_curr[self.fixed]
        fr = self.free
        info = {"newton": 0, "ls": 0, "pcg": 0, "status": "ok", "stalled": False}
        t0 = time.perf_counter()
        g = self.potential_gradient(x, xt)
        t0 = self._tick("gradient", t0)
        g0 = np.linalg.norm(g[fr])
        stalled = False
        for _ in range(self.newton_max_iter):
            gn = np.linalg.norm(g[fr])
            if gn <= self.newton_rtol * g0 or gn == 0.0 or stalled:
                info["stalled"] = bool(stalled and gn > self.newton_rtol * g0)
                break
            t0 = time.perf_counter()
            H = self.potential_hessian(x)
            t0 = self._tick("hessian", t0)
            dx = np.zeros_like(x)
            if self.linsolve == "spsolve":
                Hff = H[fr][:, fr]
                dx[fr] = spla.spsolve(Hff.tocsc(), -g[fr])
            else:
                dx, it = block_jacobi_pcg(H, -g, fr, rtol=self.pcg_rtol)
                info["pcg"] += it
            t0 = self._tick("solve", t0)
            slope = g[fr] @ dx[fr]
            t = 1.0
            accepted = False
            for _ls in range(100):
                info["ls"] += 1
                if self.energy_delta(x, x + t * dx, xt) <= 0.01 * t * slope + 1e-12:
                    accepted = True
                    break
                t *= 0.5
            t0 = self._tick("line_search", t0)
            info["newton"] += 1
            if not accepted:
                info["status"] = "line_search_failed"
                break
            x = x + t * dx
            # an update below the position resolution (4 eps max |x|) cannot improve x further
            stalled = np.abs(t * dx[fr]).max() <= 4.0 * np.finfo(float).eps * np.abs(x[fr]).max()
            g = self.potential_gradient(x, xt)
            self._tick("gradient", t0)
        else:
            info["status"] = "newton_max_iter"
        return (x, info) if return_info else x

    def rollout(self, steps, x0=None, return_info=False):
        """Run `steps` BE steps from rest (x_prev = x_curr = X); returns the list of states."""
        start = self.X.reshape(-1).copy() if x0 is None else x0.copy()
        hist = [start, start.copy(), start.copy(), start.copy()]
        states, infos = [], []
        for n in range(steps):
            x_next, info = self.step(hist[0], hist[1], (n + 1) * self.h, return_info=True, x_prev2=hist[2],
                                     x_prev3=hist[3])
            states.append(x_next)
            infos.append(info)
            hist = [x_next] + hist[:3]
        return (states, infos) if return_info else states
