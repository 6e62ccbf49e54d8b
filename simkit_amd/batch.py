"""Batched backward-Euler steps over many independent tet meshes (sk_batch_*).

Every problem gets the step simkit's backward_euler / newton_solver /
backtracking_line_search would give it alone (the device simulator's
mixed-precision variant: FP32 Hessian, PSD projection and inner PCG with
FP64 residual replacement, FP64 gradient and line-search energies), but the
whole batch advances with one launch per Newton phase: mesh-wide element and
assembly kernels over the union mesh, persistent per-problem PCG and
line-search kernels, and a hipGraph replay per batch Newton iteration.

``cantilever_batch`` builds the 512-beam workload: 4x4-cell cross-sections,
lengths U{8..64} cells, E log-uniform in [1e5, 1e7], nu U[0.30, 0.48], the
x = 0 face pinned, gravity.
"""

import ctypes
import json
import os
import time
import weakref

import numpy as np
import torch

from . import _lib
from ._device import dev, ptr, require_gpu, stream
from ._mesh import mesh_for_rest
from .gravity_force import gravity_force


def _destroy(h):
    try:
        _lib.load().sk_batch_destroy(h)
    except Exception:
        pass


class BatchSim:
    def __init__(self, meshes, mu, lam, rho=1000.0, h=1.0 / 60.0, material="snh", fixed_vertices=None, gravity=0.0,
                 psd=True, newton_rtol=1e-6, newton_max_iter=100, pcg_rtol=1e-8, pcg_max_iter=100000,
                 inner_rtol=1e-2, ls_alpha=0.01, ls_beta=0.5, ls_max_iter=100, ls_threshold=1e-12, block=0, lanes=0,
                 lds=0, use_graph=True, coarse_rows=0, coarse_period=3):
        """meshes: list of (X (n,3), T (t,4)); mu, lam: per problem (scalar each) or per union element;
        fixed_vertices: list (per problem, local vertex ids) or None.
        coarse_rows > 0: two-level per-problem PCG with rigid-motion aggregates of that many
        consecutive vertices (a multiple of 64); 0: block-Jacobi. coarse_period: re-form the
        coarse operators at the start of every coarse_period-th step (0: once, at the first step).
        The operators carry over between calls: with a period above 1 a second rollout starts from
        the last ones formed, so its positions agree with the first to the PCG tolerance rather
        than bitwise (a fresh BatchSim given the same calls reproduces them bitwise)."""
        require_gpu()
        P = len(meshes)
        Xs = [np.asarray(X, np.float64) for X, _ in meshes]
        Ts = [np.asarray(T, np.int64) for _, T in meshes]
        nvs = np.array([X.shape[0] for X in Xs])
        nts = np.array([T.shape[0] for T in Ts])
        self.vstart = np.concatenate([[0], np.cumsum(nvs)]).astype(np.int64)
        self.tstart = np.concatenate([[0], np.cumsum(nts)]).astype(np.int64)
        X = np.concatenate(Xs)
        T = np.concatenate([T + self.vstart[p] for p, T in enumerate(Ts)])
        self.P, self.X, self.T = P, X, T
        self.nv, self.nt = X.shape[0], T.shape[0]
        self.h = float(h)
        if material not in _lib.MATERIALS:
            raise ValueError("Unknown material type: " + str(material))
        self.material = _lib.MATERIALS[material]
        self.mesh, vol = mesh_for_rest(X, T)
        self.vol = vol

        def per_elem(v):
            v = np.asarray(v, np.float64).reshape(-1)
            if v.size == 1:
                return np.full(self.nt, v[0])
            if v.size == P:
                return np.repeat(v, nts)
            assert v.size == self.nt
            return v

        self.mu = dev(per_elem(mu))
        self.lam = None if lam is None else dev(per_elem(lam))
        self.mass = self.mesh.lumped_mass(vol, per_elem(rho)).repeat_interleave(3).contiguous()
        fixed = np.zeros(3 * self.nv, np.uint8)
        if fixed_vertices is not None:
            for p, fv in enumerate(fixed_vertices):
                fv = np.asarray(fv, np.int64).reshape(-1) + self.vstart[p]
                for c in range(3):
                    fixed[3 * fv + c] = 1
        self.fixed = dev(fixed, torch.uint8)
        self.f_ext = None
        if gravity:
            rho_np = per_elem(rho)
            f = np.concatenate([gravity_force(Xs[p], Ts[p], gravity, rho_np[self.tstart[p]]).reshape(-1)
                                for p in range(P)])
            self.f_ext = dev(f)
        prm = _lib.SkBatchParams()
        prm.material = self.material
        prm.psd = _lib.SK_PSD_PROJ if psd else _lib.SK_PSD_NONE
        prm.h = self.h
        prm.newton_rtol = newton_rtol
        prm.newton_max_iter = newton_max_iter
        prm.pcg_rtol = pcg_rtol
        prm.pcg_max_iter = pcg_max_iter
        prm.inner_rtol = inner_rtol
        prm.ls_alpha, prm.ls_beta, prm.ls_max_iter, prm.ls_threshold = ls_alpha, ls_beta, ls_max_iter, ls_threshold
        prm.block, prm.lanes, prm.lds, prm.use_graph = block, lanes, lds, 1 if use_graph else 0
        self.params = prm
        Th = np.ascontiguousarray(T, dtype=np.int32)
        h_ = ctypes.c_void_p()
        _lib.call("sk_batch_create", self.mesh.handle, P, self.vstart.ctypes.data_as(ctypes.c_void_p),
                  self.tstart.ctypes.data_as(ctypes.c_void_p), Th.ctypes.data_as(ctypes.c_void_p),
                  ctypes.byref(prm), ptr(self.mu), ptr(self.lam), ptr(self.vol), ptr(self.mass), ptr(self.f_ext),
                  ptr(self.fixed), ctypes.byref(h_))
        self._h = h_
        self._fin = weakref.finalize(self, _destroy, h_.value)
        self.X_dev = dev(X.reshape(-1))
        self._info = (_lib.SkBatchInfo * P)()
        self.tuned = None
        self.coarse_rows = int(coarse_rows)
        if self.coarse_rows > 0:
            Xc = np.ascontiguousarray(X, dtype=np.float64)
            _lib.call("sk_batch_set_coarse", self._h, Xc.ctypes.data_as(ctypes.c_void_p), self.coarse_rows,
                      int(coarse_period))

    def problem_slice(self, p):
        return slice(3 * int(self.vstart[p]), 3 * int(self.vstart[p + 1]))

    def step(self, x_curr, x_prev, x_bc=None, out=None, want_info=True):
        x_next = torch.empty_like(x_curr) if out is None else out
        rounds = ctypes.c_int()
        _lib.call("sk_batch_step", self._h, ptr(x_curr), ptr(x_prev), ptr(x_bc), ptr(x_next),
                  self._info if want_info else None, ctypes.byref(rounds), stream())
        infos = [self._info[p].as_dict() for p in range(self.P)] if want_info else None
        return x_next, infos, rounds.value

    def rollout(self, steps, x0=None, keep=None, want_info=True):
        x_prev = (self.X_dev if x0 is None else dev(x0)).clone()
        x_curr = x_prev.clone()
        kept, infos, rounds = {}, [], []
        for n in range(steps):
            x_next, info, r = self.step(x_curr, x_prev, want_info=want_info)
            infos.append(info)
            rounds.append(r)
            if keep is None or (n + 1) in keep:
                kept[n + 1] = x_next.cpu().numpy()
            x_prev, x_curr = x_curr, x_next
        return kept, infos, rounds

    def autotune(self, reps=3, cache=None):
        """Pick (block, lanes, lds) of the per-problem kernels by timing them on this batch's first
        Newton system from rest (lds: the PCG vectors live in LDS, tried when the largest problem
        fits); with a cache path, reuse / record the choice keyed by the batch's shape."""
        key = f"P{self.P}_nv{self.nv}_nt{self.nt}_mat{self.material}" + (
            f"_coarse{self.coarse_rows}" if self.coarse_rows else "")
        if cache and os.path.exists(cache):
            with open(cache) as fh:
                table = json.load(fh)
            if key in table:
                t = table[key]
                self._apply_choice(t["block"], t["lanes"], t.get("lds", 0))
                self.tuned = t
                return self.tuned
        b, l, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        ms = (ctypes.c_double * 16)()
        x = self.X_dev
        _lib.call("sk_batch_autotune", self._h, ptr(x), ptr(x), None, int(reps), ctypes.byref(b), ctypes.byref(l),
                  ctypes.byref(d), ms, stream())
        cands = [(ld, nt, ln) for ld in (0, 1) for nt in (128, 256, 512, 1024) for ln in (1, 4)]
        self.tuned = {"block": b.value, "lanes": l.value, "lds": d.value,
                      "ms": {f"{'lds' if ld else 'global'}_{nt}x{ln}": ms[i]
                             for i, (ld, nt, ln) in enumerate(cands) if ms[i] >= 0}}
        if cache:
            table = {}
            if os.path.exists(cache):
                with open(cache) as fh:
                    table = json.load(fh)
            table[key] = self.tuned
            os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
            with open(cache, "w") as fh:
                json.dump(table, fh, indent=1)
        return self.tuned

    def _apply_choice(self, block, lanes, lds):
        _lib.call("sk_batch_set_shape", self._h, int(block), int(lanes), int(lds))


def cantilever_batch(n=512, seed=0, cell=0.025, L_range=(8, 64), E_range=(1e5, 1e7), nu_range=(0.30, 0.48),
                     rho=1000.0, **kw):
    """The batched cantilever workload: returns (BatchSim, description dict)."""
    from .mesh import beam
    from .ympr_to_lame import ympr_to_lame
    rng = np.random.default_rng(seed)
    Ls = rng.integers(L_range[0], L_range[1] + 1, n)
    E = np.exp(rng.uniform(np.log(E_range[0]), np.log(E_range[1]), n))
    nu = rng.uniform(nu_range[0], nu_range[1], n)
    mu, lam = ympr_to_lame(E, nu)
    meshes, fixed = [], []
    for L in Ls:
        X, T = beam(int(L), 4, cell)
        meshes.append((X, T))
        fixed.append(np.where(X[:, 0] == 0.0)[0])
    sim = BatchSim(meshes, mu, lam, rho=rho, fixed_vertices=fixed, gravity=-9.8, **kw)
    desc = {"problems": int(n), "cell": cell, "lengths": Ls.tolist(), "E": E.tolist(), "nu": nu.tolist(),
            "vertices": int(sim.nv), "tets": int(sim.nt), "dof": int(3 * sim.nv)}
    return sim, desc


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Batched cantilever beams: 30 BE steps of the whole batch.")
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--shape", default=None,
                    help="block,lanes,lds of the per-problem kernels (e.g. 512,1,1) instead of autotuning")
    ap.add_argument("--compare-serial", type=int, default=0,
                    help="also advance the first K beams one at a time with the single-mesh simulator")
    ap.add_argument("--coarse-rows", type=int, default=0,
                    help="two-level per-problem PCG with aggregates of this many consecutive vertices (0: off)")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    t0 = time.perf_counter()
    sim, desc = cantilever_batch(a.n, use_graph=not a.no_graph, coarse_rows=a.coarse_rows)
    meshes_desc = desc
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    print(f"batch: {desc['problems']} beams, {desc['vertices']} vertices, {desc['tets']} tets ({t_setup:.1f}s setup)",
          flush=True)
    tuned = None
    if a.shape:
        block, lanes, lds = (int(v) for v in a.shape.split(","))
        sim._apply_choice(block, lanes, lds)
        tuned = {"block": block, "lanes": lanes, "lds": lds}
    elif not a.no_autotune:
        t1 = time.perf_counter()
        tuned = sim.autotune()
        print(f"autotune ({time.perf_counter() - t1:.1f}s): {tuned}", flush=True)
    x = sim.X_dev.clone()
    sim.step(x, x)                                       # warm-up (also captures the graph)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    kept, infos, rounds = sim.rollout(a.steps, keep={a.steps})
    torch.cuda.synchronize()
    dt = time.perf_counter() - t1
    flat = [i for step in infos for i in step]
    res = {"workload": "512 cantilever beams" if a.n == 512 else f"{a.n} cantilever beams", **desc,
           "steps": a.steps, "seconds": dt, "batch_steps_per_s": a.steps / dt,
           "beam_steps_per_s": a.steps * a.n / dt, "graph": not a.no_graph, "autotune": tuned,
           "coarse_rows": a.coarse_rows,
           "newton_rounds_per_step": float(np.mean(rounds)),
           "newton_per_beam_step": float(np.mean([i["newton_iters"] for i in flat])),
           "pcg_per_beam_step": float(np.mean([i["pcg_iters"] for i in flat])),
           "max_newton": int(max(i["newton_iters"] for i in flat)),
           "statuses": sorted(set(i["status"] for i in flat))}
    if a.compare_serial:
        from .mesh import beam
        from .sim import BackwardEulerSim
        from .ympr_to_lame import ympr_to_lame
        tot, nsteps = 0.0, 0
        for p in range(a.compare_serial):
            X, T = beam(int(meshes_desc["lengths"][p]), 4, meshes_desc["cell"])
            mu, lam = ympr_to_lame(meshes_desc["E"][p], meshes_desc["nu"][p])
            s1 = BackwardEulerSim(X, T, mu, lam, rho=1000.0, fixed_vertices=np.where(X[:, 0] == 0.0)[0],
                                  gravity=-9.8, precision="mixed")
            s1.rollout(1, keep=set())
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            s1.rollout(a.steps, keep=set())
            torch.cuda.synchronize()
            tot += time.perf_counter() - t2
            nsteps += a.steps
        res["serial_sample_beams"] = a.compare_serial
        res["serial_beam_steps_per_s"] = nsteps / tot
        res["batch_speedup_vs_serial"] = res["beam_steps_per_s"] / res["serial_beam_steps_per_s"]
    res.pop("lengths"), res.pop("E"), res.pop("nu")
    print(json.dumps(res), flush=True)
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(res, fh, indent=1)
    return res


if __name__ == "__main__":
    main()
