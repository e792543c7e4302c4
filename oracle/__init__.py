"""CPU oracle of the brick-map hot path - TEST INFRASTRUCTURE ONLY.

A plain-C restatement (px_oracle.c) of the reference's WGSL: march(),
traceSolid(), the particle segment DDA, the shade() shadow-ray setup and the
carve / pane-flood / maint chain, plus a textbook breadth-first labeller for
the islands.  Only the tests, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py use it, always as the checker of the HIP library.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

HIT_DTYPE = np.dtype([("key", "<u4"), ("t", "<f4"), ("steps", "<u4"), ("face", "u1"), ("mat", "u1"),
                      ("value_hi", "<u2")])
RAY_DTYPE = np.dtype([("ro", "<f4", 3), ("rd", "<f4", 3), ("max_t", "<f4"), ("flags", "<u4")])
PARTICLE_DTYPE = np.dtype([("pos", "<f4", 3), ("value", "<u4"), ("vel", "<f4", 3), ("life", "<f4")])
PARTICLE_DEPOSIT = 0x20000  # DEPOSIT flag of a particle's value (shared.wgsl:33-42)
# counter words of the particle oracle: C_SPAWNED, C_PCURSOR, C_WRITES, alive count
P_SPAWNED, P_PCURSOR, P_WRITES, P_ALIVE = range(4)

_lib = None


def default_threads() -> int:
    """Host threads of a parallel oracle call: every core, at most 16 (the
    CPU share of one GPU's host)."""
    return max(1, min(os.cpu_count() or 1, 16))


class EditParams(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("node_budget", C.c_uint32), ("vs", C.c_float), ("frame_seed", C.c_uint32)]


class OWorld(C.Structure):
    _fields_ = [("dx", C.c_int32), ("dy", C.c_int32), ("dz", C.c_int32), ("grid", C.c_void_p),
                ("pool", C.c_void_p), ("super_grid", C.c_void_p), ("mats", C.c_void_p)]


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        from .build_oracle import build
        build()
    _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


def load_variant(path: str, defines) -> C.CDLL:
    """A copy of the oracle built with -D `defines` at `path` (the tests plant
    known mistakes this way and check that a pin catches each)."""
    from .build_oracle import build
    build(out=path, defines=defines)
    return _bind(C.CDLL(path))


def _bind(lib):
    P = C.c_void_p
    W = C.POINTER(OWorld)
    lib.o_march.argtypes = [W, P, P, C.c_float, C.c_int, P, P]
    lib.o_trace_solid.argtypes = [W, P, P, C.c_float, P]
    lib.o_segment.argtypes = [W, P, P, P]
    lib.o_trace_batch.argtypes = [W, C.c_int, P, C.c_uint32, P, C.c_int]
    lib.o_frame.argtypes = [W, P, P, C.c_uint32, P, C.c_uint32, P, P, P, C.c_int]
    lib.o_island_pass.restype = C.c_int
    lib.o_island_pass.argtypes = [C.c_int32, C.c_int32, C.c_int32, P, P, P, P, C.c_uint32, P, P, C.c_uint32, P, P]
    lib.o_islands.restype = C.c_int
    lib.o_islands.argtypes = [C.c_int32, C.c_int32, C.c_int32, P, P, P, P, C.c_uint32, P]
    lib.o_hashu.restype = C.c_uint32
    lib.o_hashu.argtypes = [C.c_uint32]
    lib.o_hash3i.restype = C.c_uint32
    lib.o_hash3i.argtypes = [C.c_int32, C.c_int32, C.c_int32]
    lib.o_rnd.restype = C.c_float
    lib.o_rnd.argtypes = [C.c_uint32]
    lib.o_removes.restype = C.c_int
    lib.o_removes.argtypes = [P, P, C.c_int32, C.c_int32, C.c_int32, C.c_uint32]
    lib.o_apply_writes.restype = C.c_int
    lib.o_apply_writes.argtypes = [C.c_int32, C.c_int32, C.c_int32, P, P, P, C.c_uint32, C.c_uint32, C.c_uint32, P]
    lib.o_resolve.restype = C.c_int
    lib.o_resolve.argtypes = [C.c_int32, C.c_int32, C.c_int32, P, P, P, P, P, C.c_uint32, P, P, P, P, P, C.c_uint32,
                              P]
    lib.o_chunk_frame.restype = C.c_int
    lib.o_chunk_frame.argtypes = [C.c_int32, C.c_int32, C.c_int32, P, P, P, P, P, P, P, C.c_float, C.c_float,
                                  C.c_uint32, P, C.c_uint32, P, P, C.c_uint32, P, P, C.c_uint32, P]
    lib.o_raster.argtypes = [C.c_int32, C.c_int32, C.c_int32, P, C.c_uint32, P, P]
    lib.o_prim_sdf.restype = C.c_double
    lib.o_prim_sdf.argtypes = [P, C.c_double, C.c_double, C.c_double]
    for name in ("o_sinf", "o_cosf", "o_expf", "o_exp2f", "o_log2f"):
        getattr(lib, name).restype = C.c_float
        getattr(lib, name).argtypes = [C.c_float]
    for name in ("o_powf", "o_atan2f"):
        getattr(lib, name).restype = C.c_float
        getattr(lib, name).argtypes = [C.c_float, C.c_float]
    lib.o_march_ex.argtypes = [W, P, P, P, C.c_float, C.c_int, P, P, P, P, P]
    lib.o_sky_color.argtypes = [P, P, P]
    lib.o_material_color.argtypes = [P, P, C.c_uint32, P, P]
    lib.o_density.argtypes = [W, P]
    lib.o_render.argtypes = [W, P, P, C.c_uint32, P, P, C.c_uint32, P, C.c_int]
    lib.o_frame_begin.argtypes = [P]
    lib.o_spawn.argtypes = [P, C.c_uint32, C.c_uint32, P, P, C.c_uint32]
    lib.o_particle_step.argtypes = [W, P, C.c_uint32, C.c_float, C.c_float, C.c_float, C.c_uint32, P, P, P]
    lib.o_edit_pass.restype = C.c_int
    lib.o_edit_pass.argtypes = [C.c_int32, C.c_int32, C.c_int32, P, P, P, C.POINTER(EditParams), P, C.c_uint32, P,
                                C.c_uint32, P, C.c_uint32, P, P, P, C.c_uint32, P, P, C.c_uint32, P, P]
    lib.o_shoot.restype = C.c_int
    lib.o_shoot.argtypes = [C.c_int32, C.c_int32, C.c_int32, P, P, P, C.POINTER(ShootParams), P, P, P, P, P, P, P, P,
                            C.c_uint32, P, P]
    return lib


class World:
    """Oracle view of a world given as grid / pool / super grid / mats arrays."""

    def __init__(self, dims, grid, pool, super_grid, mats):
        self.grid = np.ascontiguousarray(grid, np.uint32)
        pool = np.ascontiguousarray(pool, np.uint16)
        if pool.size == 0:
            pool = np.zeros(512, np.uint16)
        self.pool = pool
        self.super_grid = np.ascontiguousarray(super_grid, np.uint32)
        self.mats = np.ascontiguousarray(mats, np.float32)
        self.dims = tuple(int(d) for d in dims)
        self.w = OWorld(self.dims[0], self.dims[1], self.dims[2], self.grid.ctypes.data, self.pool.ctypes.data,
                        self.super_grid.ctypes.data, self.mats.ctypes.data)

    @classmethod
    def from_scene(cls, s):
        return cls(s.dims, s.grid, s.pool, s.super_grid, s.mats)

    def march(self, ro, rd, max_t=1e9, shadow=False):
        ro = np.asarray(ro, np.float32)
        rd = np.asarray(rd, np.float32)
        out = np.zeros(1, HIT_DTYPE)
        trans = np.zeros(3, np.float32)
        load().o_march(C.byref(self.w), ro.ctypes.data, rd.ctypes.data, float(np.float32(max_t)), int(shadow),
                       out.ctypes.data, trans.ctypes.data)
        return out[0], trans

    def trace(self, kind: int, rays: np.ndarray, threads: int = 0) -> np.ndarray:
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        out = np.zeros(rays.shape[0], HIT_DTYPE)
        load().o_trace_batch(C.byref(self.w), kind, rays.ctypes.data, rays.shape[0], out.ctypes.data,
                             threads or default_threads())
        return out

    def frame(self, u52, lights, rays, threads: int = 0):
        u52 = np.ascontiguousarray(u52, np.float32)
        lights = np.ascontiguousarray(lights, np.float32)
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        n = rays.shape[0]
        count = int(lights[0]) if lights.size else 0
        count = max(0, min(count, (lights.size // 4 - 1) // 2 if lights.size >= 4 else 0))
        slots = 1 + count
        prim = np.zeros(n, HIT_DTYPE)
        sh = np.zeros(n * slots, HIT_DTYPE)
        nc = C.c_uint32(0)
        load().o_frame(C.byref(self.w), u52.ctypes.data, lights.ctypes.data, lights.size // 4, rays.ctypes.data, n,
                       prim.ctypes.data, sh.ctypes.data, C.byref(nc), threads or default_threads())
        return prim, sh, int(nc.value)


    # -- shading (px_shade.c) ------------------------------------------------
    def density(self) -> np.ndarray:
        """AO density texture: one byte per brick cell."""
        bd = [d // 8 for d in self.dims]
        out = np.zeros(bd[0] * bd[1] * bd[2], np.uint8)
        load().o_density(C.byref(self.w), out.ctypes.data)
        return out

    def render(self, u52, lights, rays, threads: int = 0, density=None) -> np.ndarray:
        """fsMain colour of every ray: (n, 4) float32 RGBA."""
        u52 = np.ascontiguousarray(u52, np.float32)
        lights = np.ascontiguousarray(lights, np.float32)
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        dens = self.density() if density is None else np.ascontiguousarray(density, np.uint8)
        out = np.zeros((rays.shape[0], 4), np.float32)
        load().o_render(C.byref(self.w), u52.ctypes.data, lights.ctypes.data, lights.size // 4, dens.ctypes.data,
                        rays.ctypes.data, rays.shape[0], out.ctypes.data, threads or default_threads())
        return out

    def march_ex(self, u52, ro, rd, max_t=1e9, shadow=False):
        """march() with (hit, trans, glow, outer, inner)."""
        u52 = np.ascontiguousarray(u52, np.float32)
        ro = np.asarray(ro, np.float32)
        rd = np.asarray(rd, np.float32)
        out = np.zeros(1, HIT_DTYPE)
        trans = np.zeros(3, np.float32)
        glow = np.zeros(3, np.float32)
        o, i = C.c_uint32(0), C.c_uint32(0)
        load().o_march_ex(C.byref(self.w), u52.ctypes.data, ro.ctypes.data, rd.ctypes.data,
                          float(np.float32(max_t)), int(shadow), out.ctypes.data, trans.ctypes.data,
                          glow.ctypes.data, C.byref(o), C.byref(i))
        return out[0], trans, glow, int(o.value), int(i.value)


def sky_color(u52, rd) -> np.ndarray:
    u52 = np.ascontiguousarray(u52, np.float32)
    rd = np.ascontiguousarray(rd, np.float32)
    out = np.zeros(3, np.float32)
    load().o_sky_color(u52.ctypes.data, rd.ctypes.data, out.ctypes.data)
    return out


def material_color(mats, u52, value: int, vp) -> np.ndarray:
    mats = np.ascontiguousarray(mats, np.float32)
    u52 = np.ascontiguousarray(u52, np.float32)
    vp = np.ascontiguousarray(vp, np.int32)
    out = np.zeros(3, np.float32)
    load().o_material_color(mats.ctypes.data, u52.ctypes.data, int(value), vp.ctypes.data, out.ctypes.data)
    return out


def fmath(name: str, *args) -> float:
    """One of the oracle's elementary functions: sin cos exp exp2 log2 pow atan2."""
    return float(getattr(load(), "o_%sf" % name)(*[float(np.float32(a)) for a in args]))


class Particles:
    """Serial restatement of the debris particle state: the ring of
    max_particles records, C_SPAWNED / C_PCURSOR / C_WRITES, the alive list
    and the write list of the last step."""

    def __init__(self, max_particles: int, spawn_budget: int = 60000, max_writes: int = 131072):
        assert max_particles & (max_particles - 1) == 0
        self.n = int(max_particles)
        self.budget = int(spawn_budget)
        self.max_writes = int(max_writes)
        self.parts = np.zeros(self.n, PARTICLE_DTYPE)
        self.ctr = np.zeros(4, np.uint32)
        self.alive = np.zeros(self.n, np.uint32)
        self.writes = np.zeros(2 * self.max_writes, np.uint32)

    def frame_begin(self) -> None:
        load().o_frame_begin(self.ctr.ctypes.data)

    def spawn(self, src: np.ndarray) -> None:
        src = np.ascontiguousarray(src, PARTICLE_DTYPE)
        load().o_spawn(self.parts.ctypes.data, self.n, self.budget, self.ctr.ctypes.data, src.ctypes.data,
                       src.shape[0])

    def step(self, world: "World", dt: float, gravity: float, vs: float) -> None:
        load().o_particle_step(C.byref(world.w), self.parts.ctypes.data, self.n, float(np.float32(dt)),
                               float(np.float32(gravity)), float(np.float32(vs)), self.max_writes,
                               self.ctr.ctypes.data, self.writes.ctypes.data, self.alive.ctypes.data)

    def alive_list(self) -> np.ndarray:
        return self.alive[:int(self.ctr[P_ALIVE])].copy()

    def write_list(self, base: int = 0) -> np.ndarray:
        n = min(int(self.ctr[P_WRITES]), self.max_writes)
        return self.writes[2 * base:2 * n].reshape(-1, 2).copy()


def island_pass(dims, vox: np.ndarray, kind: np.ndarray, mats, impacts, cap: int = 1 << 20):
    """Runs the edit chain on copies; returns (vox_after, kind_after, keys, counts, n, stats)."""
    vox = np.array(vox, np.uint16, copy=True, order="C")
    kind = np.array(kind, np.uint8, copy=True, order="C")
    mats = np.ascontiguousarray(mats, np.float32)
    imp = np.ascontiguousarray(impacts, np.float32).reshape(-1, 16)
    keys = np.zeros(cap, np.uint32)
    cnt = np.zeros(cap, np.uint32)
    n = C.c_uint32(0)
    stats = np.zeros(5, np.uint64)
    rc = load().o_island_pass(dims[0], dims[1], dims[2], vox.ctypes.data, kind.ctypes.data, mats.ctypes.data,
                              imp.ctypes.data, imp.shape[0], keys.ctypes.data, cnt.ctypes.data, cap, C.byref(n),
                              stats.ctypes.data)
    if rc != 0:
        raise MemoryError("oracle island pass failed")
    m = min(cap, n.value)
    return vox, kind, keys[:m].copy(), cnt[:m].copy(), int(n.value), stats


class EditResult:
    """Outcome of the oracle's edit pass: vox / kind after it, the labels
    (keys, counts; n = their total), the spawn() records in call order
    (spawn_calls counts them all) and the stats words of o_edit_pass."""

    def __init__(self, vox, kind, keys, counts, n, spawns, spawn_calls, stats):
        self.vox, self.kind, self.keys, self.counts, self.n = vox, kind, keys, counts, n
        self.spawns, self.spawn_calls, self.stats = spawns, spawn_calls, stats


def edit_pass(dims, vox: np.ndarray, kind: np.ndarray, mats, impacts, mode: int = 0,
              node_budget: int = 0xFFFFFFFF, vs: float = 1.0 / 80.0, frame: float = 0.0, seed_keys=None,
              cap: int = 1 << 20, spawn_cap: int = 1 << 22, lib=None, flood=None, dirty=None) -> EditResult:
    """impactJobs / alloc / carve with carveEffects / pane flood with
    shardEffects / maint / island search on copies (o_edit_pass).  flood:
    (key, id) entries a shot pushed; dirty: one byte per cell, the cells its
    takes touched (both from shoot())."""
    vox = np.array(vox, np.uint16, copy=True, order="C")
    kind = np.array(kind, np.uint8, copy=True, order="C")
    mats = np.ascontiguousarray(mats, np.float32)
    imp = np.ascontiguousarray(impacts, np.float32).reshape(-1, 16)
    sk = np.ascontiguousarray(np.zeros(0, np.uint32) if seed_keys is None else seed_keys, np.uint32)
    fl = np.ascontiguousarray(np.zeros((0, 2), np.uint32) if flood is None else flood, np.uint32).reshape(-1, 2)
    dirty_p = None
    if dirty is not None:
        dirty = np.ascontiguousarray(dirty, np.uint8)
        assert dirty.size == kind.size
        dirty_p = dirty.ctypes.data
    keys = np.zeros(cap, np.uint32)
    cnt = np.zeros(cap, np.uint32)
    sp = np.zeros(spawn_cap, PARTICLE_DTYPE)
    n, ns = C.c_uint32(0), C.c_uint32(0)
    stats = np.zeros(8, np.uint64)
    prm = EditParams(mode, min(int(node_budget), 0xFFFFFFFF), float(np.float32(vs)), frame_seed(frame))
    lib = lib or load()
    rc = lib.o_edit_pass(dims[0], dims[1], dims[2], vox.ctypes.data, kind.ctypes.data, mats.ctypes.data,
                         C.byref(prm), imp.ctypes.data, imp.shape[0], sk.ctypes.data, sk.size, fl.ctypes.data,
                         fl.shape[0], dirty_p, keys.ctypes.data, cnt.ctypes.data, cap, C.byref(n), sp.ctypes.data,
                         spawn_cap, C.byref(ns), stats.ctypes.data)
    if rc != 0:
        raise MemoryError("oracle edit pass failed")
    m = min(cap, n.value)
    return EditResult(vox, kind, keys[:m].copy(), cnt[:m].copy(), int(n.value), sp[:min(spawn_cap, ns.value)].copy(),
                      int(ns.value), stats)


class ShootParams(C.Structure):
    _fields_ = [("fire", C.c_float * 4), ("fire2", C.c_float * 2), ("time_x", C.c_float), ("vs", C.c_float),
                ("eye", C.c_float * 3), ("look", C.c_float * 3), ("frame_seed", C.c_uint32),
                ("max_impacts", C.c_uint32)]


class ShootResult:
    """Outcome of o_shoot: the world after the takes, the impact queue
    (impacts[:min(n_impacts, 64)], n_impacts = C_IMPACTS), the flood entries
    and the dirtied cells for edit_pass, the flash ring and its counter, the
    spawn records in call order and [destroyed, bill, pellets that hit]."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def shoot(dims, vox: np.ndarray, kind: np.ndarray, mats, u52, cam64, impacts=None, n_impacts: int = 0,
          flashes=None, flash_ctr: int = 0, max_impacts: int = 64, lib=None) -> ShootResult:
    """shoot() (sim.wgsl:371-403) on copies of the dense world: the fire /
    fire2 / time words of the uniforms u52 and eye / look of the camera
    block cam64."""
    vox = np.array(vox, np.uint16, copy=True, order="C")
    kind = np.array(kind, np.uint8, copy=True, order="C")
    mats = np.ascontiguousarray(mats, np.float32)
    u = np.asarray(u52, np.float32)
    cam = np.asarray(cam64, np.float32)
    imp = np.zeros((64, 16), np.float32) if impacts is None else np.array(impacts, np.float32).reshape(64, 16)
    fl = np.zeros((64, 2), np.uint32)
    dirty = np.zeros(kind.size, np.uint8)
    fla = np.zeros((8, 8), np.float32) if flashes is None else np.array(flashes, np.float32).reshape(8, 8)
    ni, nf, fc, ns = C.c_uint32(n_impacts), C.c_uint32(0), C.c_uint32(flash_ctr), C.c_uint32(0)
    spawn_cap = 1 << 16
    sp = np.zeros(spawn_cap, PARTICLE_DTYPE)
    stats = np.zeros(3, np.uint64)
    prm = ShootParams((C.c_float * 4)(*u[40:44]), (C.c_float * 2)(*u[44:46]), float(u[28]), float(u[3]),
                      (C.c_float * 3)(*cam[32:35]), (C.c_float * 3)(*cam[36:39]), frame_seed(float(u[29])),
                      max_impacts)
    lib = lib or load()
    lib.o_shoot(dims[0], dims[1], dims[2], vox.ctypes.data, kind.ctypes.data, mats.ctypes.data, C.byref(prm),
                imp.ctypes.data, C.byref(ni), fl.ctypes.data, C.byref(nf), dirty.ctypes.data, fla.ctypes.data,
                C.byref(fc), sp.ctypes.data, spawn_cap, C.byref(ns), stats.ctypes.data)
    return ShootResult(vox=vox, kind=kind, impacts=imp, n_impacts=int(ni.value), flood=fl[:min(64, nf.value)].copy(),
                       n_flood=int(nf.value), dirty=dirty, flashes=fla, flash_ctr=int(fc.value),
                       spawns=sp[:min(spawn_cap, ns.value)].copy(), spawn_calls=int(ns.value), stats=stats)


def apply_writes(dims, vox: np.ndarray, kind: np.ndarray, writes: np.ndarray, max_alloc: int = 65535,
                 free_top: int = 1 << 30):
    """Voxel writes + maint on copies; returns (vox_after, kind_after, stats)."""
    vox = np.array(vox, np.uint16, copy=True, order="C")
    kind = np.array(kind, np.uint8, copy=True, order="C")
    w = np.ascontiguousarray(writes, np.uint32).reshape(-1, 2)
    stats = np.zeros(4, np.uint64)
    rc = load().o_apply_writes(dims[0], dims[1], dims[2], vox.ctypes.data, kind.ctypes.data, w.ctypes.data,
                               w.shape[0], max_alloc, free_top, stats.ctypes.data)
    if rc != 0:
        raise MemoryError("oracle writes failed")
    return vox, kind, stats


class ResolveParams(C.Structure):
    _fields_ = [("max_chunks", C.c_uint32), ("max_arena", C.c_uint32), ("component_budget", C.c_uint32),
                ("max_alloc", C.c_uint32), ("free_top", C.c_uint32), ("vs", C.c_float), ("frame_seed", C.c_uint32)]


# counter words of the chunk oracle
(R_C_CHUNKS, R_C_ARENA, R_S_ISLANDS, R_S_BUDGET, R_S_BILL, R_S_DESTROYED, R_ALLOCS, R_ALLOC_FAIL, R_S_FALLING,
 R_C_WRITES, R_C_IMPACTS) = range(11)
R_WORDS = 12


def frame_seed(frame: float) -> int:
    """frameSeed() of sim.wgsl:69: u32(U.time.y) * 0x9E3779B9."""
    f = float(np.float32(frame))
    u = 0 if not f > 0 else min(int(f), 0xFFFFFFFF)
    return (u * 0x9E3779B9) & 0xFFFFFFFF


class Chunks:
    """Serial restatement of island extraction and the falling chunks: the
    chunk table, the arena, the counters, and the write / impact lists the
    chunk stages push to."""

    def __init__(self, dims, mats, max_chunks=256, max_arena=1 << 20, component_budget=400000, max_alloc=65535,
                 max_writes=131072, max_impacts=64):
        self.dims = tuple(int(d) for d in dims)
        self.mats = np.ascontiguousarray(mats, np.float32)
        self.max_chunks, self.max_arena = max_chunks, max_arena
        self.component_budget, self.max_alloc = component_budget, max_alloc
        self.max_writes, self.max_impacts = max_writes, max_impacts
        self.chunks = np.zeros(max_chunks * 16, np.uint32)
        self.arena = np.zeros(2 * max_arena, np.uint32)
        self.ctr = np.zeros(R_WORDS, np.uint32)
        self.writes = np.zeros(2 * max_writes, np.uint32)
        self.impacts = np.zeros(max_impacts * 16, np.float32)

    def params(self, vs, frame, free_top):
        return ResolveParams(self.max_chunks, self.max_arena, self.component_budget, self.max_alloc,
                             min(int(free_top), 0xFFFFFFFF), float(np.float32(vs)), frame_seed(frame))

    def resolve(self, vox, kind, keys, counts, vs, frame, free_top=1 << 31, spawn_cap=1 << 22):
        """Returns (vox_after, kind_after, spawn records in order)."""
        vox = np.array(vox, np.uint16, copy=True, order="C")
        kind = np.array(kind, np.uint8, copy=True, order="C")
        keys = np.ascontiguousarray(keys, np.uint32)
        counts = np.ascontiguousarray(counts, np.uint32)
        sp = np.zeros(spawn_cap, PARTICLE_DTYPE)
        ns = C.c_uint32(0)
        prm = self.params(vs, frame, free_top)
        d = self.dims
        rc = load().o_resolve(d[0], d[1], d[2], vox.ctypes.data, kind.ctypes.data, self.mats.ctypes.data,
                              keys.ctypes.data, counts.ctypes.data, keys.size, C.byref(prm), self.chunks.ctypes.data,
                              self.arena.ctypes.data, self.ctr.ctypes.data, sp.ctypes.data, spawn_cap, C.byref(ns))
        if rc != 0:
            raise MemoryError("oracle resolve failed")
        return vox, kind, sp[:min(spawn_cap, ns.value)].copy()

    def frame(self, vox, kind, dt, gravity, vs, frame, spawn_cap=1 << 22):
        """One frame of the chunk stages; returns spawn records in order."""
        vox = np.ascontiguousarray(vox, np.uint16)
        kind = np.ascontiguousarray(kind, np.uint8)
        sp = np.zeros(spawn_cap, PARTICLE_DTYPE)
        ns = C.c_uint32(0)
        seeds = np.zeros(1 << 20, np.uint32)
        nseed = C.c_uint32(0)
        prm = self.params(vs, frame, 0)
        d = self.dims
        rc = load().o_chunk_frame(d[0], d[1], d[2], vox.ctypes.data, kind.ctypes.data, self.mats.ctypes.data,
                                  self.chunks.ctypes.data, self.arena.ctypes.data, self.ctr.ctypes.data,
                                  C.byref(prm), float(np.float32(dt)), float(np.float32(gravity)), self.max_writes,
                                  self.writes.ctypes.data, self.max_impacts, self.impacts.ctypes.data,
                                  sp.ctypes.data, spawn_cap, C.byref(ns), seeds.ctypes.data, seeds.size,
                                  C.byref(nseed))
        if rc != 0:
            raise MemoryError("oracle chunk frame failed")
        self.land_seeds = seeds[:min(seeds.size, nseed.value)].copy()
        return sp[:min(spawn_cap, ns.value)].copy()

    def write_list(self):
        n = min(int(self.ctr[R_C_WRITES]), self.max_writes)
        return self.writes[:2 * n].reshape(-1, 2).copy()

    def impact_list(self):
        n = min(int(self.ctr[R_C_IMPACTS]), self.max_impacts)
        return self.impacts[:16 * n].reshape(-1, 16).copy()


def raster(dims, prims: np.ndarray):
    """apply() of every primitive in order on an empty world; returns
    (dense voxels shaped (dz, dy, dx), per-cell kinds)."""
    dx, dy, dz = (int(d) for d in dims)
    vox = np.zeros((dz, dy, dx), np.uint16)
    kind = np.zeros((dx // 8) * (dy // 8) * (dz // 8), np.uint8)
    pr = np.ascontiguousarray(prims)
    load().o_raster(dx, dy, dz, pr.ctypes.data, pr.shape[0], vox.ctypes.data, kind.ctypes.data)
    return vox, kind


def prim_sdf(prim, x, y, z) -> float:
    pr = np.ascontiguousarray(np.atleast_1d(prim))
    return float(load().o_prim_sdf(pr.ctypes.data, x, y, z))


def islands(dims, vox: np.ndarray, mats, cap: int = 1 << 20):
    vox = np.ascontiguousarray(vox, np.uint16)
    mats = np.ascontiguousarray(mats, np.float32)
    keys = np.zeros(cap, np.uint32)
    cnt = np.zeros(cap, np.uint32)
    n = C.c_uint32(0)
    rc = load().o_islands(dims[0], dims[1], dims[2], vox.ctypes.data, mats.ctypes.data, keys.ctypes.data,
                          cnt.ctypes.data, cap, C.byref(n))
    if rc != 0:
        raise MemoryError("oracle islands failed")
    m = min(cap, n.value)
    return keys[:m].copy(), cnt[:m].copy(), int(n.value)


def hashu(x: int) -> int:
    return int(load().o_hashu(x & 0xFFFFFFFF))


def hash3i(x: int, y: int, z: int) -> int:
    return int(load().o_hash3i(x, y, z))


def rnd(seed: int) -> float:
    return float(load().o_rnd(seed & 0xFFFFFFFF))


def removes(impact16, mats, x, y, z, v) -> bool:
    imp = np.ascontiguousarray(impact16, np.float32)
    mats = np.ascontiguousarray(mats, np.float32)
    return bool(load().o_removes(imp.ctypes.data, mats.ctypes.data, x, y, z, v))
