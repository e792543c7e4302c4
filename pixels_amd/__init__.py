"""pixels_amd - MI355X (gfx950) implementation of the sparse brick-map voxel
hot path of Crunchy000/pixels: nested DDA traversal and the destruction edit
chain, behind a C ABI (include/pixels.h) with this Python mirror of the
reference's GpuWorld / Simulation API.
"""
from . import _native
from ._native import (CTRL, CTRL_NAMES, FACE_MISS, FACE_NOT_CAST, FACE_ZERO, HIT_DTYPE, PARTICLE_DTYPE, PX_NONE,
                      RAY_DTYPE,
                      RAY_MIXED, RAY_PRIMARY, RAY_SEGMENT, RAY_SHADOW, RAY_SHOOT, PxError, PxLimits,
                      default_limits)
from .scene import LARGE, SMALL, SMALL_FLOOR, Scene
from .world import GpuWorld, Simulation

__all__ = [
    "CTRL", "CTRL_NAMES", "FACE_MISS", "FACE_NOT_CAST", "FACE_ZERO", "HIT_DTYPE", "PARTICLE_DTYPE", "PX_NONE",
    "RAY_DTYPE",
    "RAY_MIXED", "RAY_PRIMARY", "RAY_SEGMENT", "RAY_SHADOW", "RAY_SHOOT", "PxError", "PxLimits", "default_limits",
    "LARGE", "SMALL", "SMALL_FLOOR", "Scene", "GpuWorld", "Simulation", "_native",
]
