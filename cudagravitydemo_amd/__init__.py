"""MI355X-native Barnes-Hut gravity engine.

The engine is libgravity.so (hand-written gfx950 HIP kernels behind the
reference's C host API); this package binds it with ctypes (_lib), mirrors
the reference's main loop (engine.Simulation) and generates the seeded
initial conditions of the benchmark configurations (ics).
"""
from ._lib import (GravityError, LibraryMissing, Node, TreeState, WalkStats, WorldBox,
                   exported_symbols, load_library)
from .engine import Simulation
from . import ics

__all__ = ["GravityError", "LibraryMissing", "Node", "TreeState", "WalkStats", "WorldBox",
           "exported_symbols", "load_library", "Simulation", "ics"]
