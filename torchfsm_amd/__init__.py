"""torchfsm_amd: MI355X-native ETD time stepping behind torchfsm's Operator / MeshGrid /
integrate API.  Hand-written HIP kernels for gfx950 in libtfsm.so, bound through a C ABI
(include/tfsm.h)."""

from .mesh import FourierMesh, MeshGrid
from .operator import *  # noqa: F401,F403

__version__ = "0.1.0"
