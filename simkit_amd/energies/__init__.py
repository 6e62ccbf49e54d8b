"""Energies on the GPU path: stable neo-Hookean, ARAP and the backward-Euler kinetic terms."""

from .arap import *  # noqa: F401,F403
from .kinetic import *  # noqa: F401,F403
from .stable_neo_hookean import *  # noqa: F401,F403
