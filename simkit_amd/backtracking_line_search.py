"""Armijo backtracking line search (backtracking_line_search.py:8-67).

Control flow around a user objective; the objective's own work (for the
energies of this package) runs on the GPU. The device-resident simulator
(simkit_amd.sim) runs the same test on device reductions instead.
"""

import numpy as np


def backtracking_line_search(f, x0, g, dx, alpha=0.01, beta=0.5, max_iter=100, threshold=1e-12):
    """Return (t, x0 + t dx, f(x0 + t dx)) for the first t in {1, beta, beta^2, ...}
    with f(x0 + t dx) <= f(x0) + alpha t g^T dx + threshold, else (0.0, x0, f(x0))."""
    assert alpha > 0 and alpha <= 0.5
    assert beta > 0 and beta < 1
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
