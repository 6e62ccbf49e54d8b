"""Write golden vectors from the reference simkit (run in the build container only).

Covers only what closed forms, invariants and finite differences cannot pin:

1. the reference's hand-written SNH element tables (energy, P, 81-entry H);
2. ``svd_rv`` sign branches on inverted F, through the ARAP element tier and
   ``rotation_gradient_F``;
3. ``psd_project`` outputs on projected SNH Hessians;
4. the assembled CSR of ``stable_neo_hookean_hessian_x`` / ``arap_hessian_x``
   on a small grid (which entries survive scipy's exact-zero dropping);
5. the hand-written element tables of the Macklin-Mueller neo-Hookean and
   fixed-corotational energies and the linear-elasticity element tier
   (``ref_materials.npz``);
6. a 20-step backward-Euler trajectory on the 10^3-cell cube (1331 vertices,
   5000 tets; E=1e6, nu=0.45, rho=1000, gravity -9.8 in y, y=0 vertices
   eliminated) produced by the reference's own ``backward_euler`` /
   ``newton_solver`` / ``backtracking_line_search`` (spsolve, step-norm stop at
   1e-10).

Inputs are generated here with a fixed seed and stored next to the outputs.
Usage: ``python tests/golden/make_reference_golden.py [element assembly
materials trajectory]`` (needs /root/reference; no argument writes all).
"""

import os
import sys
import types

import numpy as np
import scipy.sparse as sps

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REF)
sys.path.insert(0, os.path.abspath(os.path.join(OUT, "..", "..")))

import simkit  # noqa: E402

# Load per-energy submodules without running simkit.energies.__init__ (the
# reference's own tests/conftest.py:19-36 does the same).
_stub = types.ModuleType("simkit.energies")
_stub.__path__ = [os.path.join(os.path.dirname(simkit.__file__), "energies")]
sys.modules["simkit.energies"] = _stub

from simkit.energies import stable_neo_hookean as snh  # noqa: E402
from simkit.energies import arap  # noqa: E402
from simkit.energies import fcr  # noqa: E402
from simkit.energies import linear_elasticity as le  # noqa: E402
from simkit.energies import macklin_mueller_neo_hookean as mmnh  # noqa: E402
from simkit.psd_project import psd_project  # noqa: E402
from simkit.rotation_gradient import rotation_gradient_F  # noqa: E402
from simkit.deformation_jacobian import deformation_jacobian  # noqa: E402
from simkit.volume import volume  # noqa: E402
from simkit.massmatrix import massmatrix  # noqa: E402
from simkit.gravity_force import gravity_force  # noqa: E402
from simkit.integrators.backward_euler import backward_euler  # noqa: E402

from oracle.mesh import unit_cube  # noqa: E402  (input generation only)


def random_F(rng, n, inverted=0):
    F = np.eye(3)[None] + 0.3 * rng.standard_normal((n, 3, 3))
    if inverted:
        F[:inverted, :, 0] *= -1.0          # reflect the first `inverted` gradients
    return F


def element_fixtures():
    rng = np.random.default_rng(20240501)
    F = random_F(rng, 64, inverted=12)
    mu = rng.uniform(1e4, 1e6, (64, 1))
    lam = mu * rng.uniform(1.0, 20.0, (64, 1))
    He = snh.stable_neo_hookean_hessian_element_F(F, mu, lam)
    np.savez_compressed(
        os.path.join(OUT, "ref_element.npz"),
        F=F, mu=mu, lam=lam,
        snh_psi=snh.stable_neo_hookean_energy_element_F(F, mu, lam),
        snh_P=snh.stable_neo_hookean_gradient_element_F(F, mu, lam),
        snh_H=He,
        snh_H_psd=psd_project(He),
        arap_psi=arap.arap_energy_element_F(F, mu),
        arap_P=arap.arap_gradient_element_F(F, mu),
        arap_H=arap.arap_hessian_element_F(F, mu),
        dRdF=rotation_gradient_F(F),
    )


def material_fixtures():
    rng = np.random.default_rng(20240502)
    F = random_F(rng, 64, inverted=12)
    mu = rng.uniform(1e4, 1e6, (64, 1))
    lam = mu * rng.uniform(1.0, 20.0, (64, 1))
    np.savez_compressed(
        os.path.join(OUT, "ref_materials.npz"),
        F=F, mu=mu, lam=lam,
        mmnh_psi=mmnh.macklin_mueller_neo_hookean_energy_element_F(F, mu, lam),
        mmnh_P=mmnh.macklin_mueller_neo_hookean_gradient_element_F(F, mu, lam),
        mmnh_H=mmnh.macklin_mueller_neo_hookean_hessian_element_F(F, mu, lam),
        fcr_psi=fcr.fcr_energy_element_F(F, mu, lam),
        fcr_P=fcr.fcr_gradient_element_F(F, mu, lam),
        fcr_H=fcr.fcr_hessian_element_F(F, mu, lam),
        linear_psi=le.linear_elasticity_energy_element_F(F, mu, lam),
        linear_P=le.linear_elasticity_gradient_element_F(F, mu, lam),
        linear_H=le.linear_elasticity_hessian_element_F(F, mu, lam),
    )


def assembly_fixtures():
    rng = np.random.default_rng(7)
    X, T = unit_cube(3)
    J = deformation_jacobian(X, T)
    vol = volume(X, T)
    t = T.shape[0]
    mu = np.full((t, 1), 3.0e5)
    lam = np.full((t, 1), 2.0e6)
    U = X + 0.05 * rng.standard_normal(X.shape)
    out = {"X": X, "T": T, "U": U, "mu": mu, "lam": lam}
    for name, x in (("rest", X), ("def", U)):
        for mat, fn in (("snh", lambda x: snh.stable_neo_hookean_hessian_x(x, J, mu, lam, vol, psd=True)),
                        ("arap", lambda x: arap.arap_hessian_x(x, J, mu, vol, psd=True))):
            Q = sps.csr_matrix(fn(x))
            Q.sort_indices()
            out[f"{mat}_{name}_indptr"] = Q.indptr
            out[f"{mat}_{name}_indices"] = Q.indices
            out[f"{mat}_{name}_data"] = Q.data
        out[f"snh_{name}_g"] = snh.stable_neo_hookean_gradient_x(x, J, mu, lam, vol)
        out[f"snh_{name}_E"] = snh.stable_neo_hookean_energy_x(x, J, mu, lam, vol)
        out[f"arap_{name}_g"] = arap.arap_gradient_x(x, J, mu, vol)
        out[f"arap_{name}_E"] = arap.arap_energy_x(x, J, mu, vol)
    np.savez_compressed(os.path.join(OUT, "ref_assembly.npz"), **out)


def trajectory_fixture(steps=20, keep=(1, 2, 5, 10, 20)):
    X, T = unit_cube(10)
    t = T.shape[0]
    E, nu, rho, h = 1e6, 0.45, 1000.0, 1.0 / 60.0
    mu_s = E / (2 * (1 + nu))
    lam_s = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = np.full((t, 1), mu_s)
    lam = np.full((t, 1), lam_s)
    J = deformation_jacobian(X, T)
    vol = volume(X, T)
    m = np.repeat(massmatrix(X, T, rho=rho).diagonal(), 3)
    f_ext = gravity_force(X, T, a=-9.8, rho=rho).reshape(-1)
    fixed = np.zeros(3 * X.shape[0], bool)
    for c in range(3):
        fixed[3 * np.where(X[:, 1] == 0.0)[0] + c] = True
    free = ~fixed
    x_fix = X.reshape(-1)[fixed]

    def expand(z):
        x = np.empty(3 * X.shape[0])
        x[free] = z.reshape(-1)
        x[fixed] = x_fix
        return x

    def energy(z):
        x = expand(z)
        return snh.stable_neo_hookean_energy_x(x.reshape(-1, 3), J, mu, lam, vol) - f_ext @ x

    def gradient(z):
        x = expand(z)
        g = snh.stable_neo_hookean_gradient_x(x.reshape(-1, 3), J, mu, lam, vol).reshape(-1) - f_ext
        return g[free].reshape(-1, 1)

    def hessian(z):
        x = expand(z)
        Q = sps.csr_matrix(snh.stable_neo_hookean_hessian_x(x.reshape(-1, 3), J, mu, lam, vol, psd=True))
        return Q[free][:, free]

    Mf = sps.diags(m[free])
    z_prev = X.reshape(-1)[free].reshape(-1, 1).copy()
    z_curr = z_prev.copy()
    kept = {}
    for n in range(1, steps + 1):
        z_next = backward_euler(z_curr, z_prev, energy, gradient, hessian, Mf, h,
                                tolerance=1e-10, max_iter=100)
        z_prev, z_curr = z_curr, z_next
        if n in keep:
            kept[f"x_{n}"] = expand(z_next)
    np.savez_compressed(os.path.join(OUT, "ref_be_cube10.npz"), steps=np.array(keep), **kept)


if __name__ == "__main__":
    jobs = {"element": element_fixtures, "assembly": assembly_fixtures, "materials": material_fixtures,
            "trajectory": trajectory_fixture}
    for name in sys.argv[1:] or list(jobs):
        jobs[name]()
    print("wrote", sorted(f for f in os.listdir(OUT) if f.startswith("ref_")))
