"""Material dispatch over the GPU energies, with the reference's material
strings ('linear-elasticity', 'arap', 'fcr', 'macklin-mueller-neo-hookean')
and its error (simkit/energies/elastic.py:86-113: ``ValueError("Unknown
material type: ...")``). Element, x, u and self-contained tiers; the reduced
(z, filtered z) and stretch (S) tiers are outside the position path and are
not provided.
"""

from .. import _lib
from ..psd_project import psd_project as _psd_project
from . import arap as _arap
from . import fcr as _fcr
from . import linear_elasticity as _le
from . import macklin_mueller_neo_hookean as _mm

_PREFIX = {
    "linear-elasticity": (_le, "linear_elasticity"),
    "arap": (_arap, "arap"),
    "fcr": (_fcr, "fcr"),
    "macklin-mueller-neo-hookean": (_mm, "macklin_mueller_neo_hookean"),
}


def _fn(material, kind):
    if material not in _PREFIX:
        raise ValueError("Unknown material type: " + str(material))
    mod, prefix = _PREFIX[material]
    return getattr(mod, f"{prefix}_{kind}")


def _call(material, kind, args, mu, lam, rest, kwargs):
    f = _fn(material, kind)
    if material == "arap":  # ARAP reads mu only
        return f(*args, mu, *rest, **kwargs)
    return f(*args, mu, lam, *rest, **kwargs)


def elastic_energy_element_F(F, mu, lam, material):
    return _call(material, "energy_element_F", (F,), mu, lam, (), {})


def elastic_gradient_element_F(F, mu, lam, material):
    return _call(material, "gradient_element_F", (F,), mu, lam, (), {})


def elastic_hessian_element_F(F, mu, lam, material, psd=True):
    Q = _call(material, "hessian_element_F", (F,), mu, lam, (), {})
    return _psd_project(Q) if psd else Q


def elastic_energy_x(X, J, mu, lam, vol, material):
    return _call(material, "energy_x", (X, J), mu, lam, (vol,), {})


def elastic_gradient_x(X, J, mu, lam, vol, material):
    return _call(material, "gradient_x", (X, J), mu, lam, (vol,), {})


def elastic_hessian_x(X, J, mu, lam, vol, material, psd=True):
    return _call(material, "hessian_x", (X, J), mu, lam, (vol,), {"psd": psd})


def elastic_energy_u(u, J, Jx_bar, mu, lam, vol, material):
    return _call(material, "energy_u", (u, J, Jx_bar), mu, lam, (vol,), {})


def elastic_gradient_u(u, J, Jx_bar, mu, lam, vol, material):
    return _call(material, "gradient_u", (u, J, Jx_bar), mu, lam, (vol,), {})


def elastic_hessian_u(u, J, Jx_bar, mu, lam, vol, material, psd=True):
    return _call(material, "hessian_u", (u, J, Jx_bar), mu, lam, (vol,), {"psd": psd})


def elastic_energy(X, T, mu, lam, material, U=None):
    return _call(material, "energy", (X, T), mu, lam, (), {"U": U})


def elastic_gradient(X, T, mu, lam, material, U=None):
    return _call(material, "gradient", (X, T), mu, lam, (), {"U": U})


def elastic_hessian(X, T, mu, lam, material, U=None, psd=True):
    return _call(material, "hessian", (X, T), mu, lam, (), {"U": U, "psd": psd})


MATERIAL_CODE = {
    "linear-elasticity": _lib.SK_LINEAR,
    "arap": _lib.SK_ARAP,
    "fcr": _lib.SK_FCR,
    "macklin-mueller-neo-hookean": _lib.SK_MMNH,
}
