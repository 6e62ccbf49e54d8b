"""Backward-Euler step by Newton on the incremental potential (integrators/backward_euler.py:27-91).

Generic closure form, as in the reference. For a whole mesh kept on the GPU
use simkit_amd.sim.BackwardEulerSim, which runs the same step device-resident.
"""

from ..energies.kinetic import be_target, kinetic_energy_be, kinetic_gradient_be, kinetic_hessian_be
from ..solvers import newton_solver


def backward_euler(x_curr, x_prev, energy_func, gradient_func, hessian_func, M, h, tolerance=1e-6, max_iter=1,
                   do_line_search=True, return_info=False):
    def energy(x):
        return energy_func(x) + kinetic_energy_be(x, x_curr, x_prev, M, h)

    def gradient(x):
        return gradient_func(x) + kinetic_gradient_be(x, x_curr, x_prev, M, h)

    def hessian(x):
        return hessian_func(x) + kinetic_hessian_be(M, h)

    x0 = be_target(x_curr, x_prev, h)
    return newton_solver(x0, energy, gradient, hessian, tolerance=tolerance, max_iter=max_iter,
                         do_line_search=do_line_search, return_info=return_info)
