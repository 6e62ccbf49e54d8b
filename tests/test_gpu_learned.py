"""Learned warm start (simkit_amd.learned) on a small driven cube: the modal
coordinates (basis M-orthonormal, lift exact on the driven face, project /
expand round trip), training on solver-generated steps by the library's
BF16 network, and held-out steps where the warm-started Newton reaches the
plain path's answer (1e-6 relative) in fewer Newton iterations per step than
the plain start x~: with the boundary lift alone, with the extrapolation of
its error added (no more iterations than the lift), and with the network."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_modal_coordinates_and_warm_start_small_cube():
    from simkit_amd import learned
    sim, c = learned.cube_problem(8, k=12, precision="mixed")
    X = c.X_np
    top = np.where(X[:, 1] == 1.0)[0]
    bottom = np.where(X[:, 1] == 0.0)[0]
    Phi = c.Phi.cpu().numpy()
    m = c.mass.cpu().numpy()
    assert np.abs(Phi.T @ (m[:, None] * Phi) - np.eye(12)).max() < 1e-9
    d = np.array([0.03, -0.02, 0.01])
    lift = c.lift(d).cpu().numpy().reshape(-1, 3)
    assert np.allclose(lift[top] - X[top], d, atol=1e-12) and np.allclose(lift[bottom], X[bottom], atol=1e-12)
    q = torch.as_tensor(np.linspace(-1, 1, 12) * 1e-3, dtype=torch.float64, device="cuda")
    x = c.add_modes(c.lift(d), q)
    assert torch.allclose(c.project(x, c.lift(d)), q, rtol=0, atol=1e-12)
    F, D = learned.generate(sim, c, episodes=6, steps=40, seed=1)
    assert F.shape == (240, 2 * 12 + 12) and D.shape == (240, 12)
    model = learned.WarmStartModel(12, hidden=(64, 64))
    losses = model.fit(F, D, epochs=60, batch=48, lr=2e-3)
    assert losses[-1] < 0.5 * losses[0]
    res = learned.evaluate(sim, c, model, episodes=2, steps=30, seed=2)
    print({k: (v["newton_per_step"], v.get("guess_g_ratio")) for k, v in res.items() if isinstance(v, dict)},
          flush=True)
    for start in ("lift", "extrapolated", "learned"):
        assert res[start]["final_position_rel_dev"] < 1e-6, start
        assert res[start]["steps"] == res["plain"]["steps"] == 60
    assert res["learned"]["guess_accepted"] >= 0.5
    # the gain: fewer Newton iterations per held-out step than from x~ (same tolerance)
    assert res["lift"]["newton_per_step"] < res["plain"]["newton_per_step"]
    assert res["extrapolated"]["newton_per_step"] <= res["lift"]["newton_per_step"]
    assert res["learned"]["newton_per_step"] < res["plain"]["newton_per_step"]
    assert res["extrapolated"]["guess_g_ratio"] < res["lift"]["guess_g_ratio"]
    assert res["learned"]["guess_g_ratio"] < 0.5
    # what the network adds over the same start without it: its prediction is one more candidate,
    # so the start taken is never of higher potential; Newton iterations within 0.1 per step
    print({"network_newton_reduction_over_extrapolated": res["network_newton_reduction_over_extrapolated"],
           "extrapolation_speedup_over_lift": res["extrapolation_speedup_over_lift"],
           "prediction_taken": res["learned"]["prediction_taken"]}, flush=True)
    assert res["learned"]["newton_per_step"] <= res["extrapolated"]["newton_per_step"] + 0.1
