"""BDF2 step by Newton on the incremental potential (integrators/bdf2.py:31-101).

Generic closure form, as in the reference: E(x) = V(x) + (9/4)/(2h^2)
|x - x~|_M^2 with the BDF2 target x~, minimized from x0 = x~. Bootstrap the
first step with backward_euler (no history yet). For a whole mesh kept on
the GPU use simkit_amd.sim.BackwardEulerSim.step_bdf2.
"""

from ..energies.kinetic import bdf2_target, kinetic_energy_bdf2, kinetic_gradient_bdf2, kinetic_hessian_bdf2
from ..solvers import newton_solver


def bdf2(x_curr, x_prev, x_prev2, x_prev3, energy_func, gradient_func, hessian_func, M, h, tolerance=1e-6,
         max_iter=1, do_line_search=True, return_info=False):
    def energy(x):
        return energy_func(x) + kinetic_energy_bdf2(x, x_curr, x_prev, x_prev2, x_prev3, M, h)

    def gradient(x):
        return gradient_func(x) + kinetic_gradient_bdf2(x, x_curr, x_prev, x_prev2, x_prev3, M, h)

    def hessian(x):
        return hessian_func(x) + kinetic_hessian_bdf2(M, h)

    x0 = bdf2_target(x_curr, x_prev, x_prev2, x_prev3, h)
    return newton_solver(x0, energy, gradient, hessian, tolerance=tolerance, max_iter=max_iter,
                         do_line_search=do_line_search, return_info=return_info)
