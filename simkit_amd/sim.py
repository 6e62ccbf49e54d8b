"""Device-resident backward-Euler simulator for tet meshes (the benchmark path).

One ``step`` is the reference's backward_euler -> newton_solver ->
backtracking_line_search chain (integrators/backward_euler.py:27-91,
solvers/newton.py:7-75, backtracking_line_search.py:8-67) for a stable
neo-Hookean or ARAP mesh, run entirely by the HIP library (sk_sim_step):
fused element passes, atomic-free CSR assembly, block-Jacobi PCG in place of
spsolve, and Armijo backtracking on device energy reductions. States stay in
device memory between steps.

Harness conventions (the reference leaves them to the caller):
  * pinned / prescribed vertices are eliminated, held at given positions;
  * Newton stops at ||g_free|| <= newton_rtol * ||g_free(x0)||;
  * the Armijo test is evaluated as an energy difference summed per element.
"""

import ctypes
import weakref

import numpy as np
import torch

from . import _lib
from ._device import dev, ptr, stream
from ._mesh import mesh_for_rest
from .gravity_force import gravity_force


_COARSE_MAX = 4096  # coarse unknowns the library accepts (sk_sim_set_coarse)


class BackwardEulerSim:
    def __init__(self, X, T, mu, lam=None, rho=1000.0, h=1.0 / 60.0, material="snh", fixed_vertices=(),
                 gravity=0.0, psd=True, newton_rtol=1e-6, newton_max_iter=100, pcg_rtol=1e-8,
                 pcg_max_iter=100000, ls_alpha=0.01, ls_beta=0.5, ls_max_iter=100, ls_threshold=1e-12,
                 timing=False, precision="fp64", inner_rtol=1e-2, max_refine=50, spmv_blocks=0,
                 update_blocks=0, preconditioner="bjacobi", coarse_vertices=216, pcg_chunk=0, renumber=True,
                 max_inverted=-1, pcg="classic", psd_solver="closed_form"):
        """precision: "fp64" (every stage FP64) or "mixed" (FP32 element Hessian, PSD projection,
        assembled values and inner PCG to inner_rtol, wrapped in FP64 iterative refinement to
        pcg_rtol; gradient, line-search energies and the polar SVD of ARAP elements with
        det F < 1e-3 stay FP64).

        preconditioner: "bjacobi" (3x3 block-Jacobi) or "two_level" (block-Jacobi plus an additive
        coarse correction over piecewise-rigid aggregates of about coarse_vertices rest vertices,
        raised as needed to keep the coarse space within the library's 4,096 unknowns; the coarse
        operator is formed from the first system solved). Both solve to pcg_rtol.
        renumber: the simulator numbers the vertices internally in the order its kernels read
        them (two-level: aggregate by aggregate, so that each aggregate is a contiguous range, and
        in the sliced-ELL order within an aggregate; block-Jacobi: the sliced-ELL row order, so
        that the SpMV's row stores are contiguous);
        states passed in and returned keep the caller's vertex order (the library permutes them in the
        step's first and last kernels, sk_sim_set_order).
        max_inverted: < 0 no limit; otherwise a step that ends with more elements of det F <= 0 reports
        status SK_ERR_INVERTED (info["inverted_elems"] always holds the count).
        pcg: "classic" (two launches per block-Jacobi PCG iteration) or "pipelined" (one launch per
        iteration: Ghysels-Vanroose recurrences on A B^-1); the two-level preconditioner is classic.
        psd_solver: "closed_form" (stable neo-Hookean's Hessian eigensystem from the SVD of F, the 9x9
        Jacobi for other densities and near-singular F) or "jacobi" (the 9x9 Jacobi for every element)."""
        if preconditioner not in ("bjacobi", "two_level"):
            raise ValueError("preconditioner must be 'bjacobi' or 'two_level'")
        X = np.asarray(X, dtype=np.float64)
        T = np.asarray(T)
        self.X = X
        self.T = T
        self.nv, self.nt = X.shape[0], T.shape[0]
        # aggregates of at least coarse_vertices vertices, grown until their 6 motions each fit the
        # coarse-space cap
        self.coarse_vertices = int(coarse_vertices)
        agg_order = None
        if preconditioner == "two_level":
            Xc = np.ascontiguousarray(X, dtype=np.float64)
            agg_order = np.empty(self.nv, np.int32)
            agg_ptr = np.empty(self.nv + 1, np.int32)
            n_agg = ctypes.c_int()
            self.coarse_vertices = max(self.coarse_vertices, int(np.ceil(6 * self.nv / _COARSE_MAX)))
            while True:
                _lib.call("sk_coarse_order", Xc.ctypes.data_as(ctypes.c_void_p), int(self.nv),
                          self.coarse_vertices, agg_order.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n_agg),
                          agg_ptr.ctypes.data_as(ctypes.c_void_p))
                if 6 * n_agg.value <= _COARSE_MAX:
                    break
                self.coarse_vertices = int(np.ceil(self.coarse_vertices * 1.25))
            agg_ptr = agg_ptr[:n_agg.value + 1]
        coarse_vertices = self.coarse_vertices
        # internal vertex order: order[t] = caller's vertex at internal position t
        self._order = self._pos = None
        if renumber:
            order = np.empty(self.nv, np.int32)
            Tc = np.ascontiguousarray(T, dtype=np.int32)
            _lib.call("sk_sell_order", Tc.ctypes.data_as(ctypes.c_void_p), int(self.nv), int(self.nt),
                      order.ctypes.data_as(ctypes.c_void_p))
            if preconditioner == "two_level":
                # aggregate by aggregate, each aggregate's vertices in the sliced-ELL order (descending
                # block count, ties by index), so that a SELL chunk's rows are mostly consecutive
                rank = np.empty(self.nv, np.int64)
                rank[order] = np.arange(self.nv)
                key = np.repeat(np.arange(len(agg_ptr) - 1, dtype=np.int64), np.diff(agg_ptr)) * self.nv
                order[:] = agg_order[np.argsort(key + rank[agg_order], kind="stable")]
            if not np.array_equal(order, np.arange(self.nv)):
                pos = np.empty_like(order)
                pos[order] = np.arange(self.nv, dtype=np.int32)
                self._order, self._pos = order, pos
        Xi = X if self._order is None else np.ascontiguousarray(X[self._order])
        Ti = T if self._order is None else self._pos[T]
        self.h = float(h)
        if material not in _lib.MATERIALS:
            raise ValueError("Unknown material type: " + str(material))
        self.material = _lib.MATERIALS[material]
        if self.material != _lib.SK_ARAP and lam is None:
            raise ValueError(f"material {material!r} needs lam")
        self.mesh, vol = mesh_for_rest(Xi, Ti)
        self.vol = vol
        m_v = self.mesh.lumped_mass(vol, rho)
        self.mass = m_v.repeat_interleave(3).contiguous()
        fixed = np.zeros(3 * self.nv, np.uint8)
        fv = np.asarray(fixed_vertices, dtype=np.int64).reshape(-1)
        if self._order is not None:
            fv = self._pos[fv]
        for c in range(3):
            fixed[3 * fv + c] = 1
        self.fixed_np = fixed.astype(bool)
        self.fixed = dev(fixed, torch.uint8)
        self.f_ext = None
        if gravity:
            self.f_ext = dev(gravity_force(Xi, Ti, gravity, rho).reshape(-1))
        self.mu = dev(np.broadcast_to(np.asarray(mu, np.float64).reshape(-1), (self.nt,)).copy())
        self.lam = None if lam is None else dev(np.broadcast_to(np.asarray(lam, np.float64).reshape(-1),
                                                                (self.nt,)).copy())
        p = _lib.SkSimParams()
        p.material = self.material
        p.psd = _lib.SK_PSD_PROJ if psd else _lib.SK_PSD_NONE
        p.h = self.h
        p.newton_rtol = newton_rtol
        p.newton_max_iter = newton_max_iter
        p.pcg_rtol = pcg_rtol
        p.pcg_max_iter = pcg_max_iter
        p.ls_alpha = ls_alpha
        p.ls_beta = ls_beta
        p.ls_max_iter = ls_max_iter
        p.ls_threshold = ls_threshold
        p.timing = 1 if timing else 0
        p.precision = {"fp64": _lib.SK_PREC_FP64, "mixed": _lib.SK_PREC_MIXED}[precision]
        p.inner_rtol = inner_rtol
        p.max_refine = max_refine
        p.spmv_blocks = spmv_blocks
        p.update_blocks = update_blocks
        p.pcg_chunk = pcg_chunk
        p.max_inverted = int(max_inverted)
        if pcg not in ("classic", "pipelined"):
            raise ValueError("pcg must be 'classic' or 'pipelined'")
        p.pcg_variant = 1 if pcg == "pipelined" else 0
        if psd_solver not in ("closed_form", "jacobi"):
            raise ValueError("psd_solver must be 'closed_form' or 'jacobi'")
        p.psd_jacobi = 1 if psd_solver == "jacobi" else 0
        self.params = p
        self.precision = precision
        h_ = ctypes.c_void_p()
        _lib.call("sk_sim_create", self.mesh.handle, ctypes.byref(p), ptr(self.mu), ptr(self.lam), ptr(self.vol),
                  ptr(self.mass), ptr(self.f_ext), ptr(self.fixed), ctypes.byref(h_))
        self._h = h_
        self._fin = weakref.finalize(self, _destroy, h_.value)
        if self._order is not None:
            # states cross the boundary in the caller's numbering; the library permutes them
            oc = np.ascontiguousarray(self._order, dtype=np.int32)
            _lib.call("sk_sim_set_order", self._h, oc.ctypes.data_as(ctypes.c_void_p))
        self.X_dev = dev(X.reshape(-1))
        self.preconditioner = preconditioner
        if preconditioner == "two_level":
            Xc = np.ascontiguousarray(Xi, dtype=np.float64)
            _lib.call("sk_sim_set_coarse", self._h, Xc.ctypes.data_as(ctypes.c_void_p), int(coarse_vertices))

    def _out(self, like, out):
        """Device buffer for the library's result (caller's numbering)."""
        return torch.empty_like(like) if out is None else out

    def coarse_info(self):
        """Aggregates, coarse unknowns, whether the coarse inverse has been formed and its build time (ms)."""
        n, nc, built, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _lib.call("sk_sim_coarse_info", self._h, ctypes.byref(n), ctypes.byref(nc), ctypes.byref(built),
                  ctypes.byref(ms))
        return {"aggregates": n.value, "coarse_unknowns": nc.value, "built": bool(built.value),
                "build_ms": ms.value, "vertices_per_aggregate": self.coarse_vertices}

    def step(self, x_curr, x_prev, x_bc=None, out=None):
        """One BE step on device tensors (3nv,). Returns (x_next, info dict)."""
        info = _lib.SkStepInfo()
        x_next = self._out(x_curr, out)
        _lib.call("sk_sim_step", self._h, ptr(x_curr), ptr(x_prev), ptr(x_bc), ptr(x_next), ctypes.byref(info),
                  stream())
        return x_next, info.as_dict()

    def step_guess(self, x_curr, x_prev, x_guess, x_bc=None, out=None):
        """BE step that starts Newton from x_guess when its incremental potential is lower than at
        x~ (info["guess_used"]), with the stopping test still relative to ||g(x~)||."""
        info = _lib.SkStepInfo()
        x_next = self._out(x_curr, out)
        _lib.call("sk_sim_step_guess", self._h, ptr(x_curr), ptr(x_prev), ptr(x_bc), ptr(x_guess), ptr(x_next),
                  ctypes.byref(info), stream())
        return x_next, info.as_dict()

    def step_guesses(self, x_curr, x_prev, guesses, x_bc=None, out=None):
        """BE step that starts Newton from the candidate of `guesses` (device tensors, in order) with the
        lowest incremental potential when that is below E(x~) (info["guess_used"]: its 1-based index, 0
        for x~), with the stopping test still relative to ||g(x~)||."""
        if not 0 <= len(guesses) <= 8:
            raise ValueError("step_guesses takes 0..8 guesses")
        info = _lib.SkStepInfo()
        x_next = self._out(x_curr, out)
        arr = (ctypes.c_void_p * max(1, len(guesses)))(*[ptr(g) for g in guesses])
        _lib.call("sk_sim_step_guesses", self._h, ptr(x_curr), ptr(x_prev), ptr(x_bc), len(guesses),
                  ctypes.cast(arr, ctypes.c_void_p), ptr(x_next), ctypes.byref(info), stream())
        return x_next, info.as_dict()

    def step_bdf2(self, x_curr, x_prev, x_prev2, x_prev3, x_bc=None, out=None):
        """One BDF2 step (c = 9/4, BDF2 target) on device tensors."""
        info = _lib.SkStepInfo()
        x_next = self._out(x_curr, out)
        _lib.call("sk_sim_step_bdf2", self._h, *[ptr(v) for v in (x_curr, x_prev, x_prev2, x_prev3, x_bc)],
                  ptr(x_next), ctypes.byref(info), stream())
        return x_next, info.as_dict()

    def rollout(self, steps, bc=None, x0=None, keep=None, integrator="be"):
        """`steps` steps from rest (x_prev = x_curr = X, or x0). bc(n, t) -> device x_bc for step n
        (t = (n+1) h) or None. integrator "be" or "bdf2" (BDF2 with the history levels before
        the start taken equal to the start state, i.e. from rest). Returns (kept states as host
        arrays {step: x}, infos)."""
        if integrator not in ("be", "bdf2"):
            raise ValueError("integrator must be 'be' or 'bdf2'")
        x0d = (self.X_dev if x0 is None else dev(x0)).clone()
        hist = [x0d, x0d.clone(), x0d.clone(), x0d.clone()]  # x_curr, x_prev, x_prev2, x_prev3
        kept, infos = {}, []
        for n in range(steps):
            x_bc = bc(n, (n + 1) * self.h) if bc is not None else None
            if integrator == "be":
                x_next, info = self.step(hist[0], hist[1], x_bc)
            else:
                x_next, info = self.step_bdf2(hist[0], hist[1], hist[2], hist[3], x_bc)
            infos.append(info)
            if keep is None or (n + 1) in keep:
                kept[n + 1] = x_next.cpu().numpy()
            hist = [x_next] + hist[:3]
        return kept, infos


def _destroy(h):
    try:
        _lib.load().sk_sim_destroy(h)
    except Exception:
        pass


def top_face_sinusoid(X, top_vertices, amplitude=0.1, period=1.0, axis=0):
    """bc(n, t) for BackwardEulerSim.rollout: top vertices at X + amplitude sin(2 pi t / period) e_axis,
    every other fixed vertex at X. Built from two device vectors (rest positions and a unit mask)
    combined by the library's axpby kernel into a fresh buffer per call."""
    Xd = dev(np.asarray(X, np.float64).reshape(-1))
    mask = np.zeros(Xd.numel())
    mask[3 * np.asarray(top_vertices) + axis] = 1.0
    maskd = dev(mask)

    def bc(n, t):
        out = torch.empty_like(Xd)
        _lib.call("sk_axpby", Xd.numel(), 1.0, ptr(Xd), float(amplitude * np.sin(2.0 * np.pi * t / period)),
                  ptr(maskd), ptr(out), stream())
        return out
    return bc
