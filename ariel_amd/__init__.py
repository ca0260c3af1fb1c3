"""MI355X-native FLIP step with the interface of Ariel's fluidCore::FlipSim.

The compute path is libariel_amd.so (hand-written HIP for gfx950) behind the C ABI
declared in include/ariel.h; this package is its Python mirror.
"""
from . import _abi, scenes  # noqa: F401
from .flipsim import (ArielError, FlipSim, mesh_sdf, read_levelset, write_particles,  # noqa: F401
                      write_vdb)
from .fileio import read_partio, read_vdb  # noqa: F401
from .batch import FlipBatch  # noqa: F401
