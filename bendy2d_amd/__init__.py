"""bendy2d_amd -- MI355X-native solver step of the bendy2d 2D softbody engine.

The product is the HIP library libbendy2d.so behind the C ABI in
include/bendy2d.h; this package is its Python binding (ctypes), the workload
builders and the build script.
"""
from ._lib import (BD_ENGINE_AUTO, BD_ENGINE_GLOBAL, BD_ENGINE_SCENE, BD_ENGINE_TILED, BD_ORDER_COLOUR,
                   BD_ORDER_PARTICLE, BendyError, lib)
from .solver import Batch, SceneImage, Server, Solver, plan_layers, polygon_circle, polygon_new

__all__ = ["Batch", "Server", "Solver", "plan_layers", "polygon_circle", "polygon_new", "lib", "BendyError",
           "BD_ENGINE_AUTO", "BD_ENGINE_GLOBAL", "BD_ENGINE_SCENE", "BD_ENGINE_TILED",
           "BD_ORDER_COLOUR", "BD_ORDER_PARTICLE"]
