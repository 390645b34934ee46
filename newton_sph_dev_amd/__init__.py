"""newton_sph_dev_amd: MI355X-native SPH particle step behind Newton's Model/State/Solver API.

Host code is Python; the step runs in libnsph.so (hand-written gfx950 HIP kernels, C ABI
declared in include/nsph.h) on PyTorch-ROCm tensors.
"""

from . import cbor
from .hashgrid import HashGrid
from .sim import Axis, Contacts, Control, GeoType, Model, ModelBuilder, ParticleFlags, ShapeFlags, State
from .particle_solvers import SolverSemiImplicit, SolverXPBD
from .recorder import RingBuffer, ViewerFile
from .solvers import SolverBase, SolverNotifyFlags, SolverSPH

__all__ = [
    "Axis",
    "Contacts",
    "Control",
    "GeoType",
    "HashGrid",
    "Model",
    "ModelBuilder",
    "ParticleFlags",
    "RingBuffer",
    "SolverBase",
    "SolverNotifyFlags",
    "SolverSPH",
    "ShapeFlags",
    "SolverSemiImplicit",
    "SolverXPBD",
    "State",
    "ViewerFile",
]
