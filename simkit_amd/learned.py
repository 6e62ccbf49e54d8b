"""Learned Newton warm start in a linear-modal subspace (no counterpart in the
reference, whose backward_euler.py:85 always starts Newton from the inertial
target x~).

A mesh whose fixed vertices are either pinned or driven by a uniform
displacement d (3-vector) of a face is described by

    x = X + L d + Phi q (+ a remainder M-orthogonal to Phi),

with Phi (3nv x k) the k lowest linear modes of the mesh with every fixed DOF
eliminated (linear_modal_analysis conventions: linear elasticity with
mu = lam = 1, lumped mass with rho = 1, K + 1e-6 M) and L (3nv x 3) the static
linear-elastic responses to a unit displacement of the driven face along
each axis (the "lift", so that X + L d meets the boundary values).

The warm start has three parts:

* the boundary lift: x~ = 2 x_curr - x_prev extrapolates the driven face to
  2 d_n - d_{n-1}, while the step prescribes d_{n+1}; the jump
  dd = d_{n+1} - 2 d_n + d_{n-1} is what dominates the gradient at x~ (a
  strain spike in the layer of elements under the face). x^ = x~ + L dd
  spreads it through the mesh with the static response; this is linear
  algebra, not learned;
* the extrapolation: the lift's error e_n = x_n - x^_n (the dynamic,
  nonlinear response it leaves out) changes smoothly from step to step, so
  the cubic through its last four values, x^ + 4 e_n - 6 e_{n-1} + 4 e_{n-2}
  - e_{n-3}, extrapolates it (the quadratic and linear extrapolations are
  kept as fallbacks); also not learned;
* the network: an MLP (simkit_amd.mlp, BF16 matrix cores) maps
  [q_curr, q_prev, 12 boundary-motion parameters] to the modal correction
  dq = q_next - q_b, where q_b are the modal coordinates of the cubic
  extrapolated start. The guess is that start plus Phi dq.

sk_sim_step_guesses starts Newton from the candidate of lowest incremental
potential when that is lower than x~'s, and Newton still stops at
||g|| <= rtol ||g(x~)||, so the answer meets the same tolerance from any
start. The learned start passes the prediction, then the cubic, quadratic and
linear extrapolations and the lift: a prediction worse than the extrapolation
is rejected in favour of it, not of x~. Rollouts run with start "plain" (x~, the
reference's start), "lift", "extrapolated" or "learned", so that each part's
share of the gain is measured apart from the others'.

The 12 parameters of a top-face motion d(t) = A sin(2 pi f t) (per axis) at
the step's end time t are [A (3), f (3), sin(2 pi f t) (3), cos(2 pi f t) (3)].
"""

import json
import os
import time

import numpy as np
import torch

from . import _lib
from ._device import dev, per_element, ptr, stream
from ._mesh import mesh_for_rest
from .linear_modal_analysis import modal_basis
from .mlp import MLP
from .sim import BackwardEulerSim


def _project(basis, weight, x, base, out):
    _lib.call("sk_dense_project", basis.shape[0], basis.shape[1], ptr(basis), ptr(weight), ptr(x), ptr(base),
              ptr(out), stream())
    return out


def _expand(basis, coef, base, out):
    _lib.call("sk_dense_expand", basis.shape[0], basis.shape[1], ptr(basis), ptr(coef), ptr(base), ptr(out),
              stream())
    return out


def _axpby(a, x, b, y, out):
    _lib.call("sk_axpby", x.numel(), float(a), ptr(x), float(b), ptr(y), ptr(out), stream())
    return out


class ModalCoordinates:
    """Phi, L and the projections q = Phi^T M (x - X - L d) for one mesh."""

    def __init__(self, X, T, pinned, driven, k=64, tol=1e-6, seed=0, lift_rtol=1e-10):
        X = np.asarray(X, np.float64)
        self.X_np = X
        self.nv = X.shape[0]
        self.k = int(k)
        pinned = np.asarray(pinned, np.int64).reshape(-1)
        driven = np.asarray(driven, np.int64).reshape(-1)
        fixed_v = np.concatenate([pinned, driven])
        mesh, vol = mesh_for_rest(X, T)
        self.mass = mesh.lumped_mass(vol, 1.0).repeat_interleave(3).contiguous()
        f = np.zeros(3 * self.nv, np.uint8)
        for c in range(3):
            f[3 * fixed_v + c] = 1
        ones = per_element(1.0, mesh.nt)
        self.X = dev(X.reshape(-1))
        t0 = time.perf_counter()
        self.Phi, self.eigenvalues, self.modal_info = modal_basis(
            mesh, self.X, ones, ones, vol, self.mass, dev(f, torch.uint8), self.k, shift=1e-6, tol=tol, seed=seed)
        torch.cuda.synchronize()
        self.t_modes = time.perf_counter() - t0
        # lift: static linear-elastic response to a unit displacement of the driven face per axis
        # (one implicit step with a negligible mass term, h = 1e4).
        t0 = time.perf_counter()
        stat = BackwardEulerSim(X, T, 1.0, 1.0, rho=1.0, h=1e4, material="linear", fixed_vertices=fixed_v,
                                psd=False, pcg_rtol=lift_rtol, newton_rtol=1e-8)
        cols = []
        self.lift_iters = []
        for a in range(3):
            xbc = X.copy()
            xbc[driven, a] += 1.0
            x1, info = stat.step(self.X, self.X, dev(xbc.reshape(-1)))
            if info["status"] != 0:
                raise RuntimeError(f"lift solve failed: {info}")
            self.lift_iters.append(int(info["pcg_iters"]))
            cols.append((x1 - self.X).cpu().numpy())
        self.L = dev(np.stack(cols, axis=1))
        torch.cuda.synchronize()
        self.t_lift = time.perf_counter() - t0
        self._d = torch.zeros(3, dtype=torch.float64, device=self.X.device)

    def lift(self, d, out=None):
        """X + L d (device 3nv), d a host 3-vector."""
        self._d.copy_(torch.as_tensor(np.asarray(d, np.float64)))
        out = torch.empty_like(self.X) if out is None else out
        return _expand(self.L, self._d, self.X, out)

    def project(self, x, base, out=None):
        """Phi^T M (x - base) (device k)."""
        out = torch.empty(self.k, dtype=torch.float64, device=x.device) if out is None else out
        return _project(self.Phi, self.mass, x, base, out)

    def add_modes(self, base, q, out=None):
        """base + Phi q."""
        out = torch.empty_like(base) if out is None else out
        return _expand(self.Phi, q, base, out)


def motion_params(A, f, t):
    """The 12 boundary-motion parameters at time t."""
    A = np.asarray(A, np.float64)
    f = np.asarray(f, np.float64)
    ph = 2.0 * np.pi * f * t
    return np.concatenate([A, f, np.sin(ph), np.cos(ph)])


def motion_d(A, f, t):
    return np.asarray(A, np.float64) * np.sin(2.0 * np.pi * np.asarray(f, np.float64) * t)


# The start the network's target dq is measured from: Rollout's cubic extrapolated start. A model
# trained against another base would add its correction to the wrong state.
TARGET_BASE = "extrapolated-cubic"

# Extrapolations of the lift's error, highest order first: order p takes
# sum_j EXTRAP[p][j] e_{n-j}, the degree-p polynomial through the last p + 1
# errors evaluated one step ahead (coefficients (-1)^j C(p+1, j+1)).
EXTRAP = {3: (4.0, -6.0, 4.0, -1.0), 2: (3.0, -3.0, 1.0), 1: (2.0, -1.0)}


class WarmStartModel:
    """The MLP and its input/output normalization."""

    def __init__(self, k, hidden=(256, 256, 256), seed=0, max_batch=256):
        self.k = int(k)
        self.dims = [2 * self.k + 12, *[int(h) for h in hidden], self.k]
        self.net = MLP(self.dims, max_batch=max_batch, seed=seed)
        dev_ = torch.device("cuda")
        self.in_mu = torch.zeros(self.dims[0], dtype=torch.float64, device=dev_)
        self.in_sd = torch.ones(self.dims[0], dtype=torch.float64, device=dev_)
        self.out_mu = torch.zeros(self.k, dtype=torch.float64, device=dev_)
        self.out_sd = torch.ones(self.k, dtype=torch.float64, device=dev_)

    def set_normalization(self, F, D):
        self.in_mu = F.mean(0)
        self.in_sd = F.std(0).clamp_min(1e-12)
        self.out_mu = D.mean(0)
        self.out_sd = D.std(0).clamp_min(1e-12)

    def fit(self, F, D, epochs=200, batch=256, lr=1e-3, lr_decay=0.98, seed=0, log=None):
        """F (n, 2k+12) features, D (n, k) targets dq, device float64; normalizes and trains."""
        self.set_normalization(F, D)
        Xn = ((F - self.in_mu) / self.in_sd).float().contiguous()
        Tn = ((D - self.out_mu) / self.out_sd).float().contiguous()
        return self.net.fit(Xn, Tn, epochs=epochs, batch=batch, lr=lr, seed=seed, lr_decay=lr_decay, log=log)

    def predict(self, feat, out=None):
        """dq (device k, float64) for one feature vector (device 2k+12, float64): normalization, the
        network and de-normalization in the library (sk_mlp_predict)."""
        out = torch.empty(self.k, dtype=torch.float64, device=feat.device) if out is None else out
        _lib.call("sk_mlp_predict", self.net._h, ptr(feat), 1, ptr(self.in_mu), ptr(self.in_sd), ptr(self.out_mu),
                  ptr(self.out_sd), ptr(out), stream())
        return out

    def state_dict(self):
        return {"dims": np.array(self.dims), "params": self.net.get_params(), "target_base": np.array(TARGET_BASE),
                "in_mu": self.in_mu.cpu().numpy(), "in_sd": self.in_sd.cpu().numpy(),
                "out_mu": self.out_mu.cpu().numpy(), "out_sd": self.out_sd.cpu().numpy()}

    def load_state_dict(self, d):
        assert list(d["dims"]) == self.dims
        self.net.set_params(d["params"])
        for name in ("in_mu", "in_sd", "out_mu", "out_sd"):
            setattr(self, name, dev(np.asarray(d[name], np.float64)))


STARTS = ("plain", "lift", "extrapolated", "learned")


class Rollout:
    """Steps a BackwardEulerSim through a driven-face sinusoid, tracking modal
    coordinates. start:
      "plain":        Newton from x~;
      "lift":         from x^ = x~ + L dd (the boundary lift) when its potential is lower;
      "extrapolated": from x^ plus the lift's error e_n = x_n - x^_n extrapolated from the last
                      four steps by the cubic through them (EXTRAP), falling back to the
                      quadratic and linear extrapolations, then to x^;
      "learned":      from the extrapolated start plus Phi dq, dq the model's modal correction,
                      falling back as "extrapolated" does.
    Each step starts from the candidate of lowest incremental potential (if below x~'s). The
    default is "learned" with a model, else "plain"."""

    def __init__(self, sim, coords, model=None, start=None):
        self.sim, self.c, self.model = sim, coords, model
        self.start = start or ("learned" if model is not None else "plain")
        if self.start not in STARTS:
            raise ValueError("start must be one of " + ", ".join(repr(s) for s in STARTS))
        if self.start == "learned" and model is None:
            raise ValueError("start='learned' needs a model")
        n3 = 3 * coords.nv
        z = lambda: torch.empty(n3, dtype=torch.float64, device="cuda")
        self.xt, self.xhat, self.guess, self.lift_next = z(), z(), z(), z()
        # the extrapolated starts, cubic first (xb is the cubic one); e holds e_n, ..., e_{n-3}
        self.xext = [z() for _ in EXTRAP]
        self.xb = self.xext[0]
        self.e = [z() for _ in range(max(EXTRAP) + 1)]
        self.qhat = torch.empty(coords.k, dtype=torch.float64, device="cuda")
        self.feat = torch.empty(2 * coords.k + 12, dtype=torch.float64, device="cuda")
        # per-step host scalars (12 motion parameters, the face jump dd, the face displacement d) in
        # one pinned buffer; each step synchronizes the stream, so the buffer is free again at the next
        self._host = torch.empty(18, dtype=torch.float64, pin_memory=True)
        self._dev = torch.empty(18, dtype=torch.float64, device="cuda")

    def run(self, A, f, steps, record=False, keep_last=False):
        c, sim = self.c, self.sim
        h = sim.h
        x_prev = c.X.clone()
        x_curr = c.X.clone()
        q_prev = torch.zeros(c.k, dtype=torch.float64, device="cuda")
        q_curr = torch.zeros_like(q_prev)
        feats, targets, infos = [], [], []
        d_prev = d_curr = motion_d(A, f, 0.0)          # from rest: x_prev = x_curr, face at d(0)
        ext = self.start in ("extrapolated", "learned") or record
        e = list(self.e)  # e[j] = e_{n-j}
        if ext:  # no error history at rest
            for v in e:
                _axpby(0.0, c.X, 0.0, c.X, v)
        for n in range(steps):
            t = (n + 1) * h
            d = motion_d(A, f, t)
            self._host[:12] = torch.from_numpy(motion_params(A, f, t))
            self._host[12:15] = torch.from_numpy(d - (2.0 * d_curr - d_prev))
            self._host[15:] = torch.from_numpy(d)
            self._dev.copy_(self._host, non_blocking=True)
            _expand(c.L, self._dev[15:], c.X, self.lift_next)  # X + L d: also the boundary values
            if self.start != "plain" or record:
                _axpby(2.0, x_curr, -1.0, x_prev, self.xt)
                _expand(c.L, self._dev[12:15], self.xt, self.xhat)  # x^ = x~ + L dd
            if ext:
                for out, p in zip(self.xext, sorted(EXTRAP, reverse=True)):
                    w = EXTRAP[p]  # out = x^ + sum_j w_j e_{n-j}
                    _axpby(1.0, self.xhat, w[0], e[0], out)
                    for j in range(1, len(w)):
                        _axpby(1.0, out, w[j], e[j], out)
            if self.start == "learned" or record:
                self.feat[:c.k].copy_(q_curr)
                self.feat[c.k:2 * c.k].copy_(q_prev)
                self.feat[2 * c.k:].copy_(self._dev[:12])
            if record:  # the network's target is the modal correction to the extrapolated start
                c.project(self.xb, self.lift_next, out=self.qhat)
            if self.start == "plain":
                x_next, info = sim.step(x_curr, x_prev, self.lift_next)
            elif self.start == "lift":
                x_next, info = sim.step_guess(x_curr, x_prev, self.xhat, self.lift_next)
            else:
                # candidates in order; the library starts from the one of lowest potential, so a
                # prediction worse than the extrapolation falls back to it, not to x~
                cands = self.xext + [self.xhat]
                if self.start == "learned":
                    c.add_modes(self.xb, self.model.predict(self.feat), out=self.guess)
                    cands = [self.guess] + cands
                x_next, info = sim.step_guesses(x_curr, x_prev, cands, self.lift_next)
            if info["status"] != 0:
                raise RuntimeError(f"step {n} failed: {info}")
            q_next = c.project(x_next, self.lift_next)
            if record:
                feats.append(self.feat.clone())
                targets.append(_axpby(1.0, q_next, -1.0, self.qhat, torch.empty_like(q_next)))
            infos.append(info)
            if ext:  # e_{n+1} = x_{n+1} - x^_{n+1}
                e = [e[-1]] + e[:-1]
                _axpby(1.0, x_next, -1.0, self.xhat, e[0])
            x_prev, x_curr = x_curr, x_next
            q_prev, q_curr = q_curr, q_next
            d_prev, d_curr = d_curr, d
        out = {"infos": infos}
        if record:
            out["F"] = torch.stack(feats)
            out["D"] = torch.stack(targets)
        if keep_last:
            out["x"] = x_curr
        return out


MODELS_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models")


def load_model(cells, k=64):
    """The committed model for the n^3 driven cube (trained by `python -m simkit_amd.learned --n <cells>
    ... --save`), or None."""
    path = os.path.join(MODELS_DIR, f"warmstart_cube{cells}.npz")
    if not os.path.exists(path):
        return None
    d = np.load(path, allow_pickle=False)
    base = str(d["target_base"]) if "target_base" in d else "lift"
    if base != TARGET_BASE:
        raise ValueError(f"{path}: trained against the {base!r} start, not {TARGET_BASE!r}; retrain it")
    model = WarmStartModel(k, hidden=tuple(int(h) for h in d["dims"][1:-1]))
    model.load_state_dict(d)
    return model


def random_motion(rng):
    return rng.uniform(0.0, 0.15, 3), rng.uniform(0.5, 3.0, 3)


def cube_problem(n, E=1e6, nu=0.45, rho=1000.0, h=1.0 / 60.0, precision="mixed", k=64, seed=0, **sim_kw):
    """n^3-cell unit cube, bottom pinned, top face driven; returns (sim, coords)."""
    from .mesh import unit_cube
    from .ympr_to_lame import ympr_to_lame
    X, T = unit_cube(n)
    bottom = np.where(X[:, 1] == 0.0)[0]
    top = np.where(X[:, 1] == 1.0)[0]
    mu, lam = ympr_to_lame(E, nu)
    sim = BackwardEulerSim(X, T, float(np.asarray(mu).reshape(-1)[0]), float(np.asarray(lam).reshape(-1)[0]),
                           rho=rho, h=h, fixed_vertices=np.concatenate([bottom, top]), precision=precision,
                           **sim_kw)
    coords = ModalCoordinates(X, T, bottom, top, k=k, seed=seed)
    return sim, coords


def generate(sim, coords, episodes, steps, seed, log=None):
    """Solver-generated training data: `episodes` random top-face sinusoids from rest."""
    rng = np.random.default_rng(seed)
    ro = Rollout(sim, coords, start="extrapolated")
    F, D = [], []
    for ep in range(episodes):
        A, f = random_motion(rng)
        out = ro.run(A, f, steps, record=True)
        F.append(out["F"])
        D.append(out["D"])
        if log:
            its = sum(i["newton_iters"] for i in out["infos"])
            log(f"episode {ep + 1}/{episodes}: A={np.round(A, 3).tolist()} f={np.round(f, 2).tolist()} "
                f"newton {its}")
    return torch.cat(F), torch.cat(D)


def evaluate(sim, coords, model, episodes, steps, seed, starts=STARTS):
    """Held-out episodes run from each start (Rollout's "plain", "lift", "extrapolated" and "learned"):
    Newton/PCG counts, wall time, which candidate each step started from, and agreement with the plain
    run."""
    rng = np.random.default_rng(seed)
    motions = [random_motion(rng) for _ in range(episodes)]
    res = {}
    finals = {}
    for name in starts:
        ro = Rollout(sim, coords, model if name == "learned" else None, start=name)
        ro.run(*motions[0], 2)                                 # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos, xs = [], []
        for A, f in motions:
            out = ro.run(A, f, steps, keep_last=True)
            infos += out["infos"]
            xs.append(out["x"].cpu().numpy())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        n = len(infos)
        res[name] = {"steps": n, "seconds": dt, "steps_per_s": n / dt,
                     "newton_per_step": sum(i["newton_iters"] for i in infos) / n,
                     "pcg_per_step": sum(i["pcg_iters"] for i in infos) / n,
                     "ls_per_step": sum(i["ls_trials"] for i in infos) / n,
                     "newton_per_step_list": [int(i["newton_iters"]) for i in infos]}
        if name != "plain":
            used = [i for i in infos if i["guess_used"]]
            res[name]["guess_accepted"] = len(used) / n
            # fraction of steps started from each candidate (index 1, 2, ...; 0: x~)
            res[name]["start_used"] = (np.bincount([int(i["guess_used"]) for i in infos]) / n).round(4).tolist()
            if name == "learned":  # the prediction itself, or one of its fallbacks
                res[name]["prediction_taken"] = sum(1 for i in infos if i["guess_used"] == 1) / n
                res[name]["fallback_taken"] = sum(1 for i in infos if i["guess_used"] >= 2) / n
            # gradient norm at the taken guess relative to the plain start x~ (geometric mean)
            res[name]["guess_g_ratio"] = float(np.exp(np.mean([np.log(i["guess_g_norm"] / i["g0_norm"])
                                                               for i in used]))) if used else None
        finals[name] = xs
    for name in starts:
        if name != "plain" and "plain" in finals:
            res[name]["final_position_rel_dev"] = float(max(np.abs(a - b).max() / np.abs(b).max()
                                                            for a, b in zip(finals[name], finals["plain"])))
    if "plain" in res and "learned" in res:
        res["speedup"] = res["learned"]["steps_per_s"] / res["plain"]["steps_per_s"]
        res["newton_reduction"] = res["plain"]["newton_per_step"] / res["learned"]["newton_per_step"]
        res["final_position_rel_dev"] = res["learned"]["final_position_rel_dev"]
    if "extrapolated" in res and "learned" in res:
        # the network's own share: learned against the same start without it
        res["network_speedup_over_extrapolated"] = (res["learned"]["steps_per_s"]
                                                    / res["extrapolated"]["steps_per_s"])
        res["network_newton_reduction_over_extrapolated"] = (res["extrapolated"]["newton_per_step"]
                                                             / res["learned"]["newton_per_step"])
    if "lift" in res and "extrapolated" in res:
        res["extrapolation_speedup_over_lift"] = res["extrapolated"]["steps_per_s"] / res["lift"]["steps_per_s"]
    return res


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Train and evaluate the learned warm start on a driven cube.")
    ap.add_argument("--n", type=int, default=32)
    ap.add_argument("--k", type=int, default=64)
    ap.add_argument("--episodes", type=int, default=100)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--eval-episodes", type=int, default=5)
    ap.add_argument("--eval-steps", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--precision", default="mixed")
    ap.add_argument("--save", default=None, help="write the trained model (npz)")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    t0 = time.perf_counter()
    sim, coords = cube_problem(a.n, precision=a.precision, k=a.k)
    log(f"modes: {coords.modal_info} ({coords.t_modes:.2f}s); lift {coords.t_lift:.2f}s pcg {coords.lift_iters}")
    tg = time.perf_counter()
    F, D = generate(sim, coords, a.episodes, a.steps, seed=1, log=log)
    tg = time.perf_counter() - tg
    log(f"generated {F.shape[0]} steps in {tg:.1f}s")
    model = WarmStartModel(a.k)
    tt = time.perf_counter()
    losses = model.fit(F, D, epochs=a.epochs, log=lambda ep, l: (ep % 20 == 19) and log(f"epoch {ep + 1} loss {l:.4e}"))
    torch.cuda.synchronize()
    tt = time.perf_counter() - tt
    log(f"trained in {tt:.1f}s; loss {losses[0]:.3e} -> {losses[-1]:.3e}")
    res = evaluate(sim, coords, model, a.eval_episodes, a.eval_steps, seed=2)
    res.update({"n": a.n, "k": a.k, "train_steps": int(F.shape[0]), "gen_seconds": tg, "train_seconds": tt,
                "epochs": a.epochs, "loss_first": losses[0], "loss_last": losses[-1],
                "modes": coords.modal_info, "modes_seconds": coords.t_modes, "lift_seconds": coords.t_lift,
                "total_seconds": time.perf_counter() - t0, "precision": a.precision})
    log(json.dumps(res))
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(res, fh, indent=1)
    if a.save:
        np.savez(a.save, **model.state_dict())
    return res


if __name__ == "__main__":
    main()
