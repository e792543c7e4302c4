"""Seeded scene builder (libpxscene.so): SDF primitives rasterised into the
reference's brick map (src/world/shapes.ts:334-414, VoxelWorld.ts:4-13).

Scene kinds:
  SMALL        512x256x512, 200 boxes/spheres, half-extents U[4,64], material U{1..16}
  SMALL_FLOOR  the same plus a 16-voxel structural floor slab (edit-chain scene)
  LARGE        1280x352x960, 2000 box/rounded box/cylinder/capsule/torus, sizes
               log-uniform 2-400 voxels, random yaw, structural floor + ceiling slabs
               16 voxels thick, 10% glass
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._native import P_DEPOSIT, P_DUST, PARTICLE_DTYPE, PRIM_DTYPE, RAY_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpxscene.so")

SMALL, SMALL_FLOOR, LARGE, LARGE_SCALED = 0, 1, 2, 3

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libpxscene.so is missing: build it with `python -m pixels_amd.build`")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    lib.pxs_build.restype = P
    lib.pxs_build.argtypes = [C.c_int, C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_int, C.c_double]
    lib.pxs_encode_dense.restype = P
    lib.pxs_encode_dense.argtypes = [C.c_int32, C.c_int32, C.c_int32, P, P]
    lib.pxs_free.argtypes = [P]
    lib.pxs_info.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]
    for n in ("pxs_grid", "pxs_pool", "pxs_super", "pxs_mats"):
        getattr(lib, n).restype = P
        getattr(lib, n).argtypes = [P]
    lib.pxs_decode.argtypes = [P, P]
    lib.pxs_pick_camera.restype = C.c_int
    lib.pxs_pick_camera.argtypes = [P, C.c_uint64, C.POINTER(C.c_double)]
    lib.pxs_camera_rays.argtypes = [C.POINTER(C.c_double), C.c_int32, C.c_int32, C.c_double, C.c_int32, P]
    lib.pxs_frame_inputs.argtypes = [P, C.c_uint64, C.c_int32, C.c_int32, C.c_float, C.c_int, C.c_double, P, P]
    lib.pxs_export_prims.restype = C.c_uint32
    lib.pxs_export_prims.argtypes = [P, P]
    _lib = lib
    return lib


class Scene:
    """A built world in the reference encoding, with numpy views of its arrays."""

    def __init__(self, handle):
        if not handle:
            raise ValueError("scene build failed (dims must be positive multiples of 8)")
        self._h = handle
        lib = _load()
        info = (C.c_int64 * 12)()
        counts = (C.c_uint64 * 2)()
        lib.pxs_info(handle, info, counts)
        self.dims = tuple(int(info[i]) for i in range(3))
        self.bricks = tuple(int(info[3 + i]) for i in range(3))
        self.supers = tuple(int(info[6 + i]) for i in range(3))
        self.pool_used = int(info[9])
        self.uniform_bricks = int(info[10])
        self.n_prims = int(info[11])
        self.solid = int(counts[0])
        self.glass = int(counts[1])
        ncell = self.bricks[0] * self.bricks[1] * self.bricks[2]
        nsup = self.supers[0] * self.supers[1] * self.supers[2]
        # numpy-owned copies: the arrays stay valid after the scene is freed
        self.grid = np.ctypeslib.as_array(C.cast(lib.pxs_grid(handle), C.POINTER(C.c_uint32)), (ncell,)).copy()
        self.super_grid = np.ctypeslib.as_array(C.cast(lib.pxs_super(handle), C.POINTER(C.c_uint32)),
                                                (nsup,)).copy()
        self.mats = np.ctypeslib.as_array(C.cast(lib.pxs_mats(handle), C.POINTER(C.c_float)), (256 * 20,)).copy()
        if self.pool_used:
            self.pool = np.ctypeslib.as_array(C.cast(lib.pxs_pool(handle), C.POINTER(C.c_uint16)),
                                              (self.pool_used * 512,)).copy()
        else:
            self.pool = np.zeros(0, np.uint16)

    @classmethod
    def build(cls, kind: int = SMALL, seed: int = 1, dims=(0, 0, 0), n_prims: int = 0, scale: float = 1.0):
        lib = _load()
        return cls(lib.pxs_build(kind, seed, dims[0], dims[1], dims[2], n_prims, scale))

    @classmethod
    def from_dense(cls, vox: np.ndarray, mats: np.ndarray | None = None):
        """vox: uint16 array shaped (dz, dy, dx) (x fastest in memory)."""
        vox = np.ascontiguousarray(vox, dtype=np.uint16)
        dz, dy, dx = vox.shape
        lib = _load()
        m = None if mats is None else np.ascontiguousarray(mats, dtype=np.float32)
        h = lib.pxs_encode_dense(dx, dy, dz, vox.ctypes.data, None if m is None else m.ctypes.data)
        s = cls(h)
        s._keep = (vox, m)
        return s

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.pxs_free(h)
            self._h = None

    @property
    def fill(self) -> float:
        return self.solid / float(self.dims[0] * self.dims[1] * self.dims[2])

    def dense(self) -> np.ndarray:
        """Decoded voxels shaped (dz, dy, dx)."""
        out = np.empty((self.dims[2], self.dims[1], self.dims[0]), np.uint16)
        _load().pxs_decode(self._h, out.ctypes.data)
        return out

    def camera(self, seed: int = 1):
        cam = (C.c_double * 5)()
        if _load().pxs_pick_camera(self._h, seed, cam) != 0:
            raise RuntimeError("no free camera position found")
        return [cam[i] for i in range(5)]

    def camera_rays(self, width: int, height: int, fov_deg: float = 90.0, seed: int = 1, tile: int = 8,
                    cam=None) -> np.ndarray:
        cam = cam if cam is not None else self.camera(seed)
        arr = (C.c_double * 5)(*cam)
        rays = np.zeros(width * height, RAY_DTYPE)
        _load().pxs_camera_rays(arr, width, height, fov_deg, tile, rays.ctypes.data)
        return rays

    def prims(self) -> np.ndarray:
        """The scene's SDF primitives in the px_prim layout (PRIM_DTYPE)."""
        n = _load().pxs_export_prims(self._h, None)
        out = np.zeros(n, PRIM_DTYPE)
        if n:
            _load().pxs_export_prims(self._h, out.ctypes.data)
        return out

    def frame_inputs(self, width: int, height: int, seed: int = 1, voxel_size: float = 1.0 / 80.0,
                     n_lights: int = 8, light_radius_vox: float = 4096.0):
        u = np.zeros(52, np.float32)
        lights = np.zeros((1 + 2 * n_lights) * 4, np.float32)
        _load().pxs_frame_inputs(self._h, seed, width, height, voxel_size, n_lights, light_radius_vox,
                                 u.ctypes.data, lights.ctypes.data)
        return u, lights


def debris(dims, n: int, seed: int = 1, speed: float = 60.0, life=(2.0, 8.0), deposit: float = 0.6,
           dust: float = 0.2, margin: int = 0) -> np.ndarray:
    """Synthetic debris records (px_particle) of the kind carveEffects spawns
    (sim.wgsl:495-509): voxel-centre positions uniform in the world box grown
    by `margin` voxels on x and z, velocities of magnitude up to `speed`
    voxels/s with an upward bias, a material id U{1..16} with a random shade
    byte, DEPOSIT / DUST flags with the given probabilities, and life in
    seconds uniform in `life`."""
    rng = np.random.default_rng(seed)
    out = np.zeros(n, PARTICLE_DTYPE)
    lo = np.array([-margin, 0, -margin])
    hi = np.array([dims[0] + margin, dims[1], dims[2] + margin])
    cell = rng.integers(lo, hi, size=(n, 3))
    out["pos"] = (cell + 0.5).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d[:, 1] = np.abs(d[:, 1]) * 0.7 + 0.3
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out["vel"] = (d * rng.uniform(0.0, speed, (n, 1))).astype(np.float32)
    mat = rng.integers(1, 17, n).astype(np.uint32)
    shade = rng.integers(0, 256, n).astype(np.uint32)
    flags = np.where(rng.random(n) < deposit, P_DEPOSIT, 0) | np.where(rng.random(n) < dust, P_DUST, 0)
    out["value"] = mat | (shade << 8) | flags.astype(np.uint32)
    out["life"] = rng.uniform(life[0], life[1], n).astype(np.float32)
    return out
