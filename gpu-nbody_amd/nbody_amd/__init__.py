"""nbody_amd -- Python host side of the MI355X Barnes-Hut engine.

Compute runs only in libnbody.so (HIP, gfx950), reached through its C ABI
(include/nbody.h). This package mirrors the reference's simulation plugin API
(AbstractNBodySimulation, Mode, UniverseGenerator) and provides seeded
initial conditions.
"""
from . import universe
from ._lib import NbError, load
from .simulation import (AbstractNBodySimulation, HipBarnesHutNBodySimulation, Mode,
                         UniverseGenerator, make_params)

__all__ = ["AbstractNBodySimulation", "HipBarnesHutNBodySimulation", "Mode", "UniverseGenerator",
           "NbError", "load", "make_params", "universe"]
