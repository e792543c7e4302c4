"""ctypes binding of libpixels_hip.so (the C ABI of include/pixels.h).

The library is loaded from this package directory only; a missing library is
an error (there is no CPU fallback for the hot path).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpixels_hip.so")

PX_OK = 0
PX_NONE = 0xFFFFFFFF
PX_UNIFORM_BIT = 0x80000000
PX_UNIFORM_FLOATS = 52
PX_MAT_FLOATS = 20
PX_CTRL_WORDS = 36
PX_ARGS_WORDS = 44
PX_CAMERA_FLOATS = 64

RAY_PRIMARY, RAY_SHADOW, RAY_SHOOT, RAY_SEGMENT, RAY_MIXED = 0, 1, 2, 3, 4
STAGE_EDIT, STAGE_PARTICLES, STAGE_FRAME_BEGIN, STAGE_WRITES, STAGE_RESOLVE, STAGE_CHUNKS = 1, 2, 4, 8, 16, 32
STAGE_SHOOT = 64
PX_MAX_FLASHES = 8  # layout.ts:24
P_SETTLED, P_DEPOSIT, P_DUST = 0x10000, 0x20000, 0x40000
FACE_ZERO, FACE_NOT_CAST, FACE_MISS = 6, 0xFE, 0xFF
FRAME_PIXEL_MAJOR, FRAME_SLOT_PLANAR = 0, 1
ISLANDS_ALL, ISLANDS_SEEDED = 0, 1

# CTRL_NAMES order, src/gpu/layout.ts:48-58
CTRL_NAMES = [
    "C_FREE_TOP", "C_IMPACTS", "C_JOBS", "C_ALLOCS", "C_ALLOC_N",
    "C_FLOOD_CUR", "C_FLOOD_NEXT", "C_FLOOD_PAR",
    "C_S_STATE", "C_S_CUR", "C_S_NEXT", "C_S_PAR", "C_NODES", "C_PENDING",
    "C_S_ITERS", "C_SEED_N", "C_CHUNKS", "C_ARENA", "C_WRITES", "C_WRITE_N",
    "C_DIRTY", "C_PCURSOR", "C_FLASH", "C_SPAWNED",
    "S_BILL", "S_DESTROYED", "S_ALLOC_FAIL", "S_FALLING", "S_SEARCHES", "S_ISLANDS",
    "S_SEED_DROP", "S_BUDGET", "S_HASH_FULL",
]
CTRL = {n: i for i, n in enumerate(CTRL_NAMES)}

# numpy views of the 32-byte ray and 16-byte hit records
RAY_DTYPE = np.dtype([("ro", "<f4", 3), ("rd", "<f4", 3), ("max_t", "<f4"), ("flags", "<u4")])
HIT_DTYPE = np.dtype([("key", "<u4"), ("t", "<f4"), ("steps", "<u4"), ("face", "u1"), ("mat", "u1"),
                      ("value_hi", "<u2")])
# 168-byte px_prim record (SDF primitive of the scene rasteriser)
PRIM_DTYPE = np.dtype([("type", "<u4"), ("mat", "<u4"), ("p", "<f8", 12), ("cy", "<f8"), ("sy", "<f8"),
                       ("lo", "<f8", 3), ("hi", "<f8", 3)])
PRIM_BOX, PRIM_ROUNDBOX, PRIM_SPHERE, PRIM_CYLINDER, PRIM_CAPSULE, PRIM_TORUS, PRIM_AABB = range(7)
# 32-byte px_particle record
PARTICLE_DTYPE = np.dtype([("pos", "<f4", 3), ("value", "<u4"), ("vel", "<f4", 3), ("life", "<f4")])


class PxLimits(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "max_dirty", "max_alloc", "max_writes", "max_impacts", "max_jobs", "max_flood", "max_seeds",
        "hash_size", "max_chunks", "max_arena", "max_lights", "max_anchors", "max_flashes",
        "spawn_budget", "component_budget", "search_iters_per_frame", "flood_iters_per_frame",
        "island_mode", "uf_walk_cap")]


class PxWorldDesc(C.Structure):
    _fields_ = [
        ("dims", C.c_uint32 * 3),
        ("grid", C.POINTER(C.c_uint32)),
        ("pool", C.POINTER(C.c_uint16)),
        ("pool_used", C.c_uint32),
        ("holes", C.POINTER(C.c_uint32)),
        ("n_holes", C.c_uint32),
        ("mats", C.POINTER(C.c_float)),
        ("pool_capacity", C.c_uint32),
        ("limits", C.POINTER(PxLimits)),
        ("voxel_size", C.c_float),
        ("device", C.c_int32),
    ]


class PxQuerySet(C.Structure):
    _fields_ = [("rays", C.c_void_p * 4), ("out", C.c_void_p * 4), ("n", C.c_uint32 * 4)]


# every exported symbol with (restype, argtypes)
_P = C.c_void_p
_U32 = C.c_uint32
_I32P = C.POINTER(C.c_int32)
_U32P = C.POINTER(C.c_uint32)
_FP = C.POINTER(C.c_float)
SIGNATURES = {
    "px_default_limits": (None, [C.POINTER(PxLimits)]),
    "px_create": (C.c_int, [C.POINTER(PxWorldDesc), C.POINTER(_P)]),
    "px_create_from_prims": (C.c_int, [C.POINTER(PxWorldDesc), _P, _U32, C.POINTER(_P)]),
    "px_destroy": (C.c_int, [_P]),
    "px_last_error": (C.c_char_p, [_P]),
    "px_set_uniforms": (C.c_int, [_P, _FP]),
    "px_set_lights": (C.c_int, [_P, _FP, _U32]),
    "px_sync": (C.c_int, [_P]),
    "px_stream": (_P, [_P]),
    "px_trace": (C.c_int, [_P, _U32, _P, _U32, _P]),
    "px_trace_device": (C.c_int, [_P, _U32, _P, _U32, _P]),
    "px_trace_frame_device": (C.c_int, [_P, _P, _U32, _P, _P, _U32P]),
    "px_trace_frame_device_noalloc": (C.c_int, [_P, _P, _U32, _P, _P, _U32P]),
    "px_set_frame_layout": (C.c_int, [_P, _U32]),
    "px_render_frame_device": (C.c_int, [_P, _P, _U32, _P]),
    "px_render": (C.c_int, [_P, _P, _U32, _P]),
    "px_read_density": (C.c_int, [_P, _P]),
    "px_frame_slots": (_U32, [_P]),
    "px_trace_batch": (C.c_int, [_P, C.POINTER(PxQuerySet)]),
    "px_set_batch_shape": (C.c_int, [_P, _U32, _U32, _U32]),
    "px_get_batch_shape": (C.c_int, [_P, _U32P]),
    "px_autotune_batch": (C.c_int, [_P, C.POINTER(PxQuerySet), _U32, _U32P, C.POINTER(C.c_float)]),
    "px_autotune_batch_perm": (C.c_int, [_P, C.POINTER(PxQuerySet), _P, _U32, _U32P, C.POINTER(C.c_float)]),
    "px_probe": (C.c_int, [_P, _I32P, _I32P, _U32P]),
    "px_probe_device": (C.c_int, [_P, _I32P, _I32P, _P]),
    "px_read_ctrl": (C.c_int, [_P, _U32P, _U32P, _FP]),
    "px_read_grid": (C.c_int, [_P, _U32P, _U32P]),
    "px_read_pool": (C.c_int, [_P, _U32, _U32, _U32P]),
    "px_world_info": (C.c_int, [_P, _U32P, _U32P, _U32P, _U32P]),
    "px_push_impacts": (C.c_int, [_P, _FP, _U32]),
    "px_island_pass": (C.c_int, [_P]),
    "px_sim_step": (C.c_int, [_P, _U32]),
    "px_run_to_quiescence": (C.c_int, [_P]),
    "px_island_labels": (C.c_int, [_P, _U32P, _U32P, _U32, _U32P]),
    "px_snapshot": (C.c_int, [_P]),
    "px_restore": (C.c_int, [_P]),
    "px_trace_perm_device": (C.c_int, [_P, _U32, _P, _P, _U32, _P]),
    "px_mlp_param_count": (_U32, []),
    "px_mlp_create": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "px_mlp_destroy": (C.c_int, [_P]),
    "px_mlp_get_params": (C.c_int, [_P, _FP]),
    "px_mlp_set_params": (C.c_int, [_P, _FP]),
    "px_mlp_features": (C.c_int, [_P, _U32, _P, _U32, _P]),
    "px_mlp_predict": (C.c_int, [_P, _P, _U32, _P]),
    "px_mlp_grad": (C.c_int, [_P, _P, _P, _U32, _FP, _FP]),
    "px_mlp_train_step": (C.c_int, [_P, _P, _P, _U32, C.c_float, _FP]),
    "px_bin_rays": (C.c_int, [_P, _P, _U32, _U32, _U32, C.c_float, C.c_float, _P]),
    "px_mlp_sort_rays": (C.c_int, [_P, _U32, _P, _U32, _U32, C.c_float, C.c_float, _U32, _P]),
    "px_mlp_trace_sorted": (C.c_int, [_P, _U32, _P, _U32, _U32, C.c_float, C.c_float, _U32, _P, _P]),
    "px_particles_init": (C.c_int, [_P, _U32]),
    "px_frame_begin": (C.c_int, [_P]),
    "px_spawn": (C.c_int, [_P, _P, _U32]),
    "px_spawn_device": (C.c_int, [_P, _P, _U32]),
    "px_particle_step": (C.c_int, [_P]),
    "px_read_particles": (C.c_int, [_P, _U32, _U32, _P]),
    "px_read_alive": (C.c_int, [_P, _U32P, _U32, _U32P]),
    "px_read_writes": (C.c_int, [_P, _U32P, _U32, _U32P]),
    "px_particle_buffers": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "px_apply_writes": (C.c_int, [_P]),
    "px_resolve_islands": (C.c_int, [_P]),
    "px_chunk_step": (C.c_int, [_P]),
    "px_read_chunks": (C.c_int, [_P, _U32P, _U32P, _U32, _U32P]),
    "px_read_device_impacts": (C.c_int, [_P, _FP, _U32, _U32P]),
    "px_set_camera": (C.c_int, [_P, _FP]),
    "px_trace_batch_perm": (C.c_int, [_P, _P, _P]),
    "px_capture_sim_step": (C.c_int, [_P, _U32, C.POINTER(_P)]),
    "px_graph_launch": (C.c_int, [_P, _P]),
    "px_graph_destroy": (C.c_int, [_P]),
    "px_read_flashes": (C.c_int, [_P, _FP]),
    "px_push_writes": (C.c_int, [_P, _P, _U32]),
    "px_profile_begin": (C.c_int, [_P]),
    "px_profile_end": (C.c_int, [_P, C.POINTER(C.c_double)]),
}
SORT_FACES = 1  # px_mlp_sort_rays flag: keys (direction face, cost bucket)


def SORT_WINDOW(log2: int) -> int:
    """px_mlp_sort_rays flag: sort each window of 2**log2 consecutive rays
    (log2 8..30) on its own, the windows in input order."""
    return (int(log2) & 31) << 8
PROF_PRIMARY, PROF_SHADOW, PROF_TRACE, PROF_EDIT, PROF_BATCH, PROF_MLP, PROF_PARTICLES = 0, 1, 2, 3, 4, 5, 6
PROF_PARTICLE_SIM, PROF_WRITES, PROF_CHUNKS = 7, 8, 9
PROF_DENSITY, PROF_RENDER_PRIMARY, PROF_RENDER_SHADOW, PROF_SHADE, PROF_REFLECT = 10, 11, 12, 13, 14
PROF_GRAPH = 15
PROF_SLOTS = 16

_lib = None


def load():
    """Load libpixels_hip.so; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libpixels_hip.so is missing: build it with `python -m pixels_amd.build` "
                           "(the hot path has no CPU fallback)")
    # torch ships its own libamdhip64.so.7; loading it first makes this
    # library bind to the same HIP runtime (one runtime per process), so torch
    # tensors and this library's kernels can share device memory and streams.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class PxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("pixels error %d: %s" % (code, msg))
        self.code = code


def check(rc: int, ctx=None) -> None:
    if rc != PX_OK:
        lib = load()
        raise PxError(rc, lib.px_last_error(ctx).decode(errors="replace"))


def ptr(a: np.ndarray, ctype=C.c_uint32):
    return a.ctypes.data_as(C.POINTER(ctype))


def default_limits() -> PxLimits:
    lim = PxLimits()
    load().px_default_limits(C.byref(lim))
    return lim
