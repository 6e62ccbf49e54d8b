"""Energies on the GPU path: stable neo-Hookean, ARAP, linear elasticity,
Macklin-Mueller neo-Hookean, fixed corotational, the material dispatcher and
the backward-Euler / BDF2 kinetic terms."""

from .arap import *  # noqa: F401,F403
from .elastic import *  # noqa: F401,F403
from .fcr import *  # noqa: F401,F403
from .kinetic import *  # noqa: F401,F403
from .linear_elasticity import *  # noqa: F401,F403
from .macklin_mueller_neo_hookean import *  # noqa: F401,F403
from .stable_neo_hookean import *  # noqa: F401,F403
