"""lettuce_amd: lettuce's lattice-Boltzmann time step on AMD Instinct MI355X.

The public names follow lettuce (``Context``, ``D2Q9``, ``BGKCollision``,
``TaylorGreenVortex``, ``Simulation`` ...).  Every time step runs through the
hand-written HIP kernels of ``liblettuce_amd.so``; build it with
``python -m lettuce_amd._build``.
"""
from ._context import *  # noqa: F401,F403
from ._stencil import *  # noqa: F401,F403
from ._unit import *  # noqa: F401,F403
from ._flow import *  # noqa: F401,F403
from ._simulation import *  # noqa: F401,F403
from .ext import *  # noqa: F401,F403
from .util import *  # noqa: F401,F403
from . import util, ext  # noqa: F401

__version__ = "0.1.0"
