"""Newton's method with backtracking (solvers/newton.py:7-75).

Same loop, defaults and stopping rule (||alpha dx|| < tolerance) as the
reference; the direct sparse solve is replaced by GPU PCG (simkit_amd.linsolve).
"""

import numpy as np

from ..backtracking_line_search import backtracking_line_search
from ..linsolve import pcg_solve


def newton_solver(x0, energy_func, gradient_func, hessian_func, tolerance=1e-6, max_iter=1, do_line_search=True,
                  return_info=False):
    x = x0.copy()
    if return_info:
        info = {"g": [], "dx": [], "alphas": [], "iters": -1}
    for i in range(max_iter):
        g = gradient_func(x)
        H = hessian_func(x)
        dx = pcg_solve(H, -np.asarray(g)).reshape(-1, 1)
        if do_line_search:
            alpha, lx, ex = backtracking_line_search(energy_func, x, g, dx)
        else:
            alpha = 1.0
        x += alpha * dx
        if return_info:
            info["g"].append(g)
            info["dx"].append(dx)
            info["alphas"].append(alpha)
            info["iters"] = i
        if np.linalg.norm(alpha * dx) < tolerance:
            break
    if return_info:
        return x, info
    return x
