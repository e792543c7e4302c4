"""Host-side mirror of the reference's GpuWorld / Simulation API over the C ABI.

  GpuWorld    src/gpu/GpuWorld.ts:17-137   (one-time upload, pool sizing,
                                            super counts, free list)
  Simulation  src/gpu/Simulation.ts:61-323 (probe, per-frame stats readback)
              plus the ray queries and the island pass of the hot path.

Every call goes to libpixels_hip.so; device buffers may be passed as torch
tensors (their data_ptr() is used) or host numpy arrays.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .scene import Scene


def _devptr(t) -> int:
    return int(t.data_ptr())


class GpuWorld:
    """Every device resource of one world (GpuWorld.ts:17-137)."""

    def __init__(self, scene: Scene | None, device: int = 0, limits: N.PxLimits | None = None, pool_capacity: int = 0,
                 voxel_size: float = 1.0 / 80.0, prims: np.ndarray | None = None, dims=None, mats=None):
        """From a host-built scene (px_create), or, with `prims` (PRIM_DTYPE),
        `dims` and `mats`, rasterised on the GPU (px_create_from_prims)."""
        lib = N.load()
        self.lib = lib
        self.scene = scene
        d = N.PxWorldDesc()
        d.dims[:] = scene.dims if scene is not None else dims
        self._mats = np.ascontiguousarray(scene.mats if scene is not None else mats, np.float32)
        d.holes = None
        d.n_holes = 0
        d.mats = N.ptr(self._mats, C.c_float)
        d.pool_capacity = pool_capacity
        self.limits = limits if limits is not None else N.default_limits()
        d.limits = C.pointer(self.limits)
        d.voxel_size = voxel_size
        d.device = device
        h = C.c_void_p()
        if prims is None:
            self._grid = np.ascontiguousarray(scene.grid, np.uint32)
            self._pool = np.ascontiguousarray(scene.pool, np.uint16)
            d.grid = N.ptr(self._grid)
            d.pool = N.ptr(self._pool, C.c_uint16) if self._pool.size else None
            d.pool_used = scene.pool_used
            N.check(lib.px_create(C.byref(d), C.byref(h)), None)
        else:
            pr = np.ascontiguousarray(prims, N.PRIM_DTYPE)
            d.grid = None
            d.pool = None
            d.pool_used = 0
            N.check(lib.px_create_from_prims(C.byref(d), pr.ctypes.data if pr.size else None, pr.shape[0],
                                             C.byref(h)), None)
        self.ctx = h
        dims = (C.c_uint32 * 3)()
        bricks = (C.c_uint32 * 3)()
        supers = (C.c_uint32 * 3)()
        cap = C.c_uint32()
        N.check(lib.px_world_info(h, dims, bricks, supers, C.byref(cap)), h)
        self.dims = tuple(dims)
        self.bricks = tuple(bricks)
        self.supers = tuple(supers)
        self.numCells = self.bricks[0] * self.bricks[1] * self.bricks[2]
        self.poolCapacity = int(cap.value)
        self.initialBricks = scene.pool_used if scene is not None else None

    def close(self) -> None:
        if getattr(self, "ctx", None):
            self.lib.px_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int) -> None:
        N.check(rc, self.ctx)

    @property
    def stream(self) -> int:
        return int(self.lib.px_stream(self.ctx) or 0)

    def after_torch(self) -> None:
        """Orders the library's stream after the work torch has queued on its
        current stream (an event wait, no host stall): device buffers that
        torch just wrote are complete before a kernel of the library reads
        them.  The library's results are read back after px_sync().  Tensors
        handed to the library must stay alive until then: torch's allocator
        does not see the library's stream, and marking them with
        record_stream would make torch record events on that stream after
        close() has destroyed it."""
        import torch
        if not torch.cuda.is_available():
            return
        ext = getattr(self, "_ext_stream", None)
        if ext is None:
            ext = self._ext_stream = torch.cuda.ExternalStream(self.stream)
        ext.wait_stream(torch.cuda.current_stream())

    def sync(self) -> None:
        self.check(self.lib.px_sync(self.ctx))

    def set_uniforms(self, u: np.ndarray) -> None:
        u = np.ascontiguousarray(u, np.float32)
        assert u.size == N.PX_UNIFORM_FLOATS
        self.check(self.lib.px_set_uniforms(self.ctx, N.ptr(u, C.c_float)))

    def set_lights(self, lights: np.ndarray) -> None:
        lights = np.ascontiguousarray(lights, np.float32)
        self.check(self.lib.px_set_lights(self.ctx, N.ptr(lights, C.c_float), lights.size // 4))

    def snapshot(self) -> None:
        self.check(self.lib.px_snapshot(self.ctx))

    def restore(self) -> None:
        self.check(self.lib.px_restore(self.ctx))

    def profile_begin(self) -> None:
        self.check(self.lib.px_profile_begin(self.ctx))

    def profile_end(self) -> dict:
        """{slot: (summed ms, launches)} of the kernels launched since profile_begin."""
        out = (C.c_double * (2 * N.PROF_SLOTS))()
        self.check(self.lib.px_profile_end(self.ctx, out))
        return {k: (out[2 * k], int(out[2 * k + 1])) for k in range(N.PROF_SLOTS) if out[2 * k + 1] > 0}

    def read_grid(self):
        g = np.empty(self.numCells, np.uint32)
        s = np.empty(self.supers[0] * self.supers[1] * self.supers[2], np.uint32)
        self.check(self.lib.px_read_grid(self.ctx, N.ptr(g), N.ptr(s)))
        return g, s

    def read_pool(self, first: int, count: int) -> np.ndarray:
        w = np.empty(count * 256, np.uint32)
        self.check(self.lib.px_read_pool(self.ctx, first, count, N.ptr(w)))
        return w


def impacts_from_hits(dims, rays: np.ndarray, hits: np.ndarray, radius, fragile_radius=None, impulse=0.0):
    """Impact records (16 floats, pushImpact layout sim.wgsl:203-211) for ray
    hits, centred as shoot() does (sim.wgsl:382-385):
    c = voxel + 0.5 - n * min(r * 0.3, 2)."""
    dx, dy, _ = dims
    ok = hits["face"] != N.FACE_MISS
    h = hits[ok]
    r = np.broadcast_to(np.asarray(radius, np.float32), hits.shape)[ok]
    fr = r if fragile_radius is None else np.broadcast_to(np.asarray(fragile_radius, np.float32), hits.shape)[ok]
    key = h["key"].astype(np.int64)
    vox = np.stack([key % dx, (key // dx) % dy, key // (dx * dy)], 1).astype(np.float32)
    n = np.zeros((h.shape[0], 3), np.float32)
    face = h["face"].astype(np.int64)
    sel = face < 6
    n[np.nonzero(sel)[0], face[sel] >> 1] = np.where(face[sel] & 1, -1.0, 1.0).astype(np.float32)
    off = np.minimum(r * np.float32(0.3), np.float32(2.0)).astype(np.float32)
    c = (vox + np.float32(0.5)) - n * off[:, None]
    out = np.zeros((h.shape[0], 16), np.float32)
    out[:, 0:3] = c
    out[:, 3] = r
    out[:, 4] = fr
    out[:, 5] = impulse
    out[:, 6:9] = n
    out[:, 9:12] = rays["rd"][ok]
    return out


class SimStats(dict):
    """Decoded counters, as SimStats of Simulation.ts:24-39."""


class Simulation:
    """Dispatch / queue API of Simulation.ts over the library."""

    def __init__(self, gw: GpuWorld):
        self.gw = gw
        self.lib = gw.lib
        self.stats = SimStats()

    # -- Simulation.probe (Simulation.ts:161-196) --------------------------
    def probe(self, origin, size) -> np.ndarray:
        o = (C.c_int32 * 3)(*origin)
        s = (C.c_int32 * 3)(*size)
        out = np.empty(int(size[0]) * int(size[1]) * int(size[2]), np.uint32)
        self.gw.check(self.lib.px_probe(self.gw.ctx, o, s, N.ptr(out)))
        return out

    def probe_device(self, origin, size, out_tensor) -> None:
        self.gw.after_torch()
        o = (C.c_int32 * 3)(*origin)
        s = (C.c_int32 * 3)(*size)
        self.gw.check(self.lib.px_probe_device(self.gw.ctx, o, s, C.c_void_p(_devptr(out_tensor))))

    # -- afterSubmit readback (Simulation.ts:289-322) ----------------------
    def read_ctrl(self):
        ctrl = np.zeros(N.PX_CTRL_WORDS, np.uint32)
        args = np.zeros(N.PX_ARGS_WORDS, np.uint32)
        cam = np.zeros(N.PX_CAMERA_FLOATS, np.float32)
        self.gw.check(self.lib.px_read_ctrl(self.gw.ctx, N.ptr(ctrl), N.ptr(args), N.ptr(cam, C.c_float)))
        return ctrl, args, cam

    def after_submit(self) -> SimStats:
        ctrl, args, cam = self.read_ctrl()
        c = N.CTRL
        self.stats = SimStats(
            bill=ctrl[c["S_BILL"]] / 10.0,
            destroyed=int(ctrl[c["S_DESTROYED"]]),
            falling=int(ctrl[c["S_FALLING"]]),
            liveBricks=self.gw.poolCapacity - int(ctrl[c["C_FREE_TOP"]]),
            allocFailures=int(ctrl[c["S_ALLOC_FAIL"]]),
            searches=int(ctrl[c["S_SEARCHES"]]),
            islands=int(ctrl[c["S_ISLANDS"]]),
            searching=int(ctrl[c["C_S_STATE"]]) != 0,
        )
        return self.stats

    # -- ray queries --------------------------------------------------------
    def trace(self, kind: int, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, N.RAY_DTYPE)
        out = np.zeros(rays.shape[0], N.HIT_DTYPE)
        self.gw.check(self.lib.px_trace(self.gw.ctx, kind, rays.ctypes.data, rays.shape[0], out.ctypes.data))
        return out

    def trace_device(self, kind: int, rays_t, n: int, out_t) -> None:
        self.gw.after_torch()
        self.gw.check(self.lib.px_trace_device(self.gw.ctx, kind, C.c_void_p(_devptr(rays_t)), n,
                                               C.c_void_p(_devptr(out_t))))

    def frame_slots(self) -> int:
        return int(self.lib.px_frame_slots(self.gw.ctx))

    def trace_frame_device(self, rays_t, n: int, prim_t, shadow_t, count: bool = False):
        self.gw.after_torch()
        nc = C.c_uint32(0)
        self.gw.check(self.lib.px_trace_frame_device(self.gw.ctx, C.c_void_p(_devptr(rays_t)), n,
                                                     C.c_void_p(_devptr(prim_t)), C.c_void_p(_devptr(shadow_t)),
                                                     C.byref(nc) if count else None))
        return int(nc.value) if count else None

    def set_frame_layout(self, slot_planar: bool) -> None:
        """Shadow outputs of the frame: pixel-major (slot s of ray i at
        i * slots + s) or slot-planar (at s * n + i)."""
        self.gw.check(self.lib.px_set_frame_layout(self.gw.ctx, N.FRAME_SLOT_PLANAR if slot_planar
                                                   else N.FRAME_PIXEL_MAJOR))

    # -- shading to an image (fsMain, raymarch.wgsl:387-450) ------------------
    def render(self, rays: np.ndarray) -> np.ndarray:
        """RGBA float colour of every ray (host buffers, blocking)."""
        rays = np.ascontiguousarray(rays, N.RAY_DTYPE)
        out = np.zeros((rays.shape[0], 4), np.float32)
        if rays.shape[0]:
            self.gw.check(self.lib.px_render(self.gw.ctx, N.ptr(rays), rays.shape[0], N.ptr(out)))
        return out

    def render_device(self, rays_t, n: int, rgba_t) -> None:
        """Device variant: rays_t holds n px_ray records, rgba_t 4 floats per ray."""
        self.gw.after_torch()
        self.gw.check(self.lib.px_render_frame_device(self.gw.ctx, C.c_void_p(_devptr(rays_t)), n,
                                                      C.c_void_p(_devptr(rgba_t))))

    def read_density(self) -> np.ndarray:
        """The AO density texture, one byte per brick cell (x fastest)."""
        bd = [d // 8 for d in self.gw.dims]
        out = np.zeros(bd[0] * bd[1] * bd[2], np.uint8)
        self.gw.check(self.lib.px_read_density(self.gw.ctx, N.ptr(out)))
        return out

    # -- batched queries (one persistent launch, replayed as a HIP graph) ----
    @staticmethod
    def query_set(rays_by_kind: dict, out_by_kind: dict) -> N.PxQuerySet:
        """{kind: (tensor of px_ray records, count)}, {kind: tensor of px_hit records}."""
        q = N.PxQuerySet()
        for k, (t, n) in rays_by_kind.items():
            q.rays[k] = _devptr(t) if n else None
            q.out[k] = _devptr(out_by_kind[k]) if n else None
            q.n[k] = n
        return q

    def trace_batch(self, q: N.PxQuerySet) -> None:
        self.gw.after_torch()
        self.gw.check(self.lib.px_trace_batch(self.gw.ctx, C.byref(q)))

    def trace_batch_perm(self, q: N.PxQuerySet, perms: dict) -> None:
        """px_trace_batch_perm: {kind: int32 tensor of source indices}."""
        arr = (C.c_void_p * 4)()
        for k, t in perms.items():
            arr[k] = _devptr(t)
        self.gw.after_torch()
        self.gw.check(self.lib.px_trace_batch_perm(self.gw.ctx, C.byref(q), C.cast(arr, C.c_void_p)))

    def set_batch_shape(self, block: int, blocks: int = 0, chunk: int = 256) -> None:
        self.gw.check(self.lib.px_set_batch_shape(self.gw.ctx, block, blocks, chunk))

    def get_batch_shape(self) -> dict:
        out = np.zeros(3, np.uint32)
        self.gw.check(self.lib.px_get_batch_shape(self.gw.ctx, N.ptr(out)))
        return dict(block=int(out[0]), blocks=int(out[1]), chunk=int(out[2]))

    def autotune_batch(self, q: N.PxQuerySet, repeats: int = 3, perms: dict | None = None):
        """px_autotune_batch, or px_autotune_batch_perm for the launch in the
        orders `perms` ({kind: int32 tensor}); leaves the best shape set."""
        out = np.zeros(3, np.uint32)
        ms = C.c_float(0.0)
        if perms is None:
            self.gw.check(self.lib.px_autotune_batch(self.gw.ctx, C.byref(q), repeats, N.ptr(out), C.byref(ms)))
        else:
            arr = (C.c_void_p * 4)()
            for k, t in perms.items():
                arr[k] = _devptr(t)
            self.gw.after_torch()
            self.gw.check(self.lib.px_autotune_batch_perm(self.gw.ctx, C.byref(q), C.cast(arr, C.c_void_p), repeats,
                                                          N.ptr(out), C.byref(ms)))
        return dict(block=int(out[0]), blocks=int(out[1]), chunk=int(out[2])), float(ms.value)

    # -- edit chain -----------------------------------------------------------
    def push_impacts(self, impacts: np.ndarray) -> None:
        imp = np.ascontiguousarray(impacts, np.float32).reshape(-1, 16)
        self.gw.check(self.lib.px_push_impacts(self.gw.ctx, N.ptr(imp, C.c_float), imp.shape[0]))

    def island_pass(self) -> None:
        self.gw.check(self.lib.px_island_pass(self.gw.ctx))

    def set_camera(self, cam: np.ndarray) -> None:
        """The Camera block the player pass owns (shared.wgsl:22-30): 64
        floats, eye (m) at [32:35], look at [36:39]; the shoot stage reads them."""
        cam = np.ascontiguousarray(cam, np.float32)
        assert cam.size == N.PX_CAMERA_FLOATS
        self.gw.check(self.lib.px_set_camera(self.gw.ctx, N.ptr(cam, C.c_float)))

    def read_flashes(self) -> np.ndarray:
        """The flash ring (sim.wgsl:395-401): PX_MAX_FLASHES x 8 floats."""
        out = np.zeros((N.PX_MAX_FLASHES, 8), np.float32)
        self.gw.check(self.lib.px_read_flashes(self.gw.ctx, N.ptr(out, C.c_float)))
        return out

    def read_device_impacts(self, cap: int = 64):
        """Impacts queued on the device and not yet taken by an island pass
        (16 floats each) and C_IMPACTS."""
        out = np.zeros((cap, 16), np.float32)
        n = C.c_uint32(0)
        self.gw.check(self.lib.px_read_device_impacts(self.gw.ctx, N.ptr(out, C.c_float), cap, C.byref(n)))
        return out[:min(cap, int(n.value))].copy(), int(n.value)

    def encode(self, stages: int = N.STAGE_EDIT) -> None:
        """Simulation.encode (Simulation.ts:200-273): the hot-path stages, non-blocking."""
        self.gw.check(self.lib.px_sim_step(self.gw.ctx, stages))

    def capture(self, stages: int) -> "FrameGraph":
        """px_capture_sim_step: the stages as a replayable HIP graph (each
        stage must have run once through encode())."""
        g = C.c_void_p()
        self.gw.check(self.lib.px_capture_sim_step(self.gw.ctx, stages, C.byref(g)))
        return FrameGraph(self, g)

    def run_to_quiescence(self) -> None:
        self.gw.check(self.lib.px_run_to_quiescence(self.gw.ctx))

    def island_labels(self, cap: int = 1 << 20):
        keys = np.zeros(cap, np.uint32)
        cnt = np.zeros(cap, np.uint32)
        n = C.c_uint32(0)
        self.gw.check(self.lib.px_island_labels(self.gw.ctx, N.ptr(keys), N.ptr(cnt), cap, C.byref(n)))
        m = min(cap, int(n.value))
        return keys[:m].copy(), cnt[:m].copy(), int(n.value)

    # -- debris particles (sim.wgsl:166-170, 363-369, 1082-1173) ----------------
    def particles_init(self, max_particles: int = 1 << 18) -> None:
        self.gw.check(self.lib.px_particles_init(self.gw.ctx, max_particles))
        self.max_particles = int(max_particles) or (1 << 18)

    def frame_begin(self) -> None:
        self.gw.check(self.lib.px_frame_begin(self.gw.ctx))

    def spawn(self, records: np.ndarray) -> None:
        rec = np.ascontiguousarray(records, N.PARTICLE_DTYPE)
        self.gw.check(self.lib.px_spawn(self.gw.ctx, rec.ctypes.data, rec.shape[0]))

    def spawn_device(self, records_t, n: int) -> None:
        self.gw.after_torch()
        self.gw.check(self.lib.px_spawn_device(self.gw.ctx, C.c_void_p(_devptr(records_t)), n))

    def particle_step(self) -> None:
        """particleSim with dt, gravity and voxel size from the uniforms."""
        self.gw.check(self.lib.px_particle_step(self.gw.ctx))

    def read_particles(self, first: int = 0, count: int | None = None) -> np.ndarray:
        count = self.max_particles - first if count is None else count
        out = np.zeros(count, N.PARTICLE_DTYPE)
        self.gw.check(self.lib.px_read_particles(self.gw.ctx, first, count, out.ctypes.data))
        return out

    def read_alive(self) -> np.ndarray:
        n = C.c_uint32(0)
        self.gw.check(self.lib.px_read_alive(self.gw.ctx, None, 0, C.byref(n)))
        out = np.zeros(max(1, int(n.value)), np.uint32)
        self.gw.check(self.lib.px_read_alive(self.gw.ctx, N.ptr(out), out.size, C.byref(n)))
        return out[:min(int(n.value), self.max_particles)].copy()

    def read_writes(self) -> np.ndarray:
        """(key, value) pairs of the frame's deposit writes, shape (n, 2)."""
        n = C.c_uint32(0)
        self.gw.check(self.lib.px_read_writes(self.gw.ctx, None, 0, C.byref(n)))
        out = np.zeros(2 * max(1, int(n.value)), np.uint32)
        self.gw.check(self.lib.px_read_writes(self.gw.ctx, N.ptr(out), int(n.value), C.byref(n)))
        return out[:2 * int(n.value)].reshape(-1, 2).copy()

    def particle_buffers(self):
        p, a, w = C.c_void_p(), C.c_void_p(), C.c_void_p()
        self.gw.check(self.lib.px_particle_buffers(self.gw.ctx, C.byref(p), C.byref(a), C.byref(w)))
        return int(p.value or 0), int(a.value or 0), int(w.value or 0)

    # -- voxel writes (sim.wgsl:1178-1203) ------------------------------------------
    def push_writes(self, pairs: np.ndarray) -> None:
        """Append (key, value) pairs to the write list (pushWrite from the host)."""
        w = np.ascontiguousarray(pairs, np.uint32).reshape(-1, 2)
        self.gw.check(self.lib.px_push_writes(self.gw.ctx, w.ctypes.data, w.shape[0]))

    def apply_writes(self) -> None:
        """writeArgs + writeRequest + alloc + writeApply + maint of the touched cells."""
        self.gw.check(self.lib.px_apply_writes(self.gw.ctx))


class FrameGraph:
    """A frame of px_sim_step stages recorded as a HIP graph."""

    def __init__(self, sim: "Simulation", handle):
        self.sim, self.handle = sim, handle

    def launch(self) -> None:
        self.sim.gw.check(self.sim.lib.px_graph_launch(self.sim.gw.ctx, self.handle))

    def close(self) -> None:
        if self.handle:
            self.sim.lib.px_graph_destroy(self.handle)
            self.handle = None
