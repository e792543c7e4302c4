"""GPU parity of the voxel-write stage (px_apply_writes: writeRequest,
allocBricks, writeApply, maint; sim.wgsl:94-146, 459-483, 1178-1203,
1219-1274) with the oracle's list-order restatement (oracle.apply_writes):
every decoded voxel, the brick encoding after maint, the super-grid counts
and the counters.  Cases: random writes with repeated keys and shared
columns (stacking) on the seeded scenes, a small max_alloc, a pool with few
free bricks, writes that pass max_writes, and the whole debris loop
(particles settle, their deposit writes become rubble voxels) over frames."""
import numpy as np
import pytest

import oracle
from pixels_amd import _native as N
from tests.test_gpu_edit import kinds_of

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)
VS = np.float32(1.0 / 80.0)
G = np.float32(9.81 / VS)


def _random_writes(scene, n, seed):
    """Writes near surfaces: random voxels, a share of repeated keys and of
    keys one or two voxels above earlier ones (same column)."""
    rng = np.random.default_rng(seed)
    dx, dy, dz = scene.dims
    x = rng.integers(0, dx, n)
    y = rng.integers(0, dy, n)
    z = rng.integers(0, dz, n)
    rep = rng.random(n) < 0.15
    src = rng.integers(0, n, n)
    x[rep], y[rep], z[rep] = x[src[rep]], y[src[rep]], z[src[rep]]
    up = (rng.random(n) < 0.1) & ~rep
    y[up] = np.minimum(y[src[up]] + rng.integers(1, 3, up.sum()), dy - 1)
    x[up], z[up] = x[src[up]], z[src[up]]
    key = (x + dx * (y + dy * z)).astype(np.uint32)
    val = (rng.integers(1, 17, n) | 0x80 | (rng.integers(0, 256, n) << 8)).astype(np.uint32)
    return np.stack([key, val], 1)


def _state(gw, sim, scene):
    vox = sim.probe((0, 0, 0), scene.dims).reshape(scene.dims[2], scene.dims[1], scene.dims[0])
    grid, sgrid = gw.read_grid()
    ctrl, _, _ = sim.read_ctrl()
    return vox, grid, sgrid, ctrl


def _check(scene, got, ovox, okind):
    vox, grid, sgrid, ctrl = got
    assert np.array_equal(vox, ovox.astype(np.uint32)), "voxels differ"
    assert np.array_equal(kinds_of(grid), okind), "brick encoding differs"
    bx, by, bz = scene.bricks
    sx, sy, sz = scene.supers
    pad = np.zeros((sz * 4, sy * 4, sx * 4), np.uint32)
    pad[:bz, :by, :bx] = grid.reshape(bz, by, bx) != 0
    assert np.array_equal(pad.reshape(sz, 4, sy, 4, sx, 4).sum((1, 3, 5)).ravel(), sgrid)


@pytest.mark.parametrize("case", ["plain", "max_alloc", "small_pool", "overflow"])
def test_writes_match_oracle(floor_scene, case):
    from pixels_amd import GpuWorld, Simulation
    s = floor_scene
    lim = N.default_limits()
    n = 60000
    pool_capacity = 0
    if case == "max_alloc":
        lim.max_alloc = 300
    if case == "overflow":
        lim.max_writes = 20000
    if case == "small_pool":
        pool_capacity = s.pool_used + 200
    gw = GpuWorld(s, limits=lim, pool_capacity=pool_capacity)
    sim = Simulation(gw)
    try:
        w = _random_writes(s, n, seed=5)
        ctrl0, _, _ = sim.read_ctrl()
        free_top = int(ctrl0[N.CTRL["C_FREE_TOP"]])
        print("case: writes %s n=%d free_top=%d" % (case, n, free_top), flush=True)
        sim.push_writes(w[: n // 2])
        sim.push_writes(w[n // 2:])
        ctrl1, _, _ = sim.read_ctrl()
        assert int(ctrl1[N.CTRL["C_WRITES"]]) == n
        sim.apply_writes()
        got = _state(gw, sim, s)
        m = min(n, int(lim.max_writes))
        ovox, okind, st = oracle.apply_writes(s.dims, s.dense(), kinds_of(s.grid), w[:m],
                                              max_alloc=int(lim.max_alloc), free_top=free_top)
        _check(s, got, ovox, okind)
        ctrl = got[3]
        assert int(ctrl[N.CTRL["C_WRITES"]]) == 0 and int(ctrl[N.CTRL["C_WRITE_N"]]) == m
        if case == "small_pool":
            assert int(ctrl[N.CTRL["S_ALLOC_FAIL"]]) > 0
        assert st[0] > 1000
        # an empty list is a no-op
        sim.apply_writes()
        again = _state(gw, sim, s)
        assert np.array_equal(again[0], got[0]) and np.array_equal(again[1], got[1])
    finally:
        gw.close()


def test_trace_sees_writes_inside_pool_bricks(floor_scene):
    """Traversal reads a per-brick mask of its non-empty 2x2x2 voxel blocks.
    Writes that fill air voxels of existing pool bricks (no brick changes
    occupancy) must reach it: rays traced before and after the writes equal
    the oracle's on the world before and after, bit for bit."""
    from pixels_amd import GpuWorld, Scene, Simulation
    s = floor_scene
    gw = GpuWorld(s)
    sim = Simulation(gw)
    try:
        rays = s.camera_rays(192, 108, 90.0, seed=11)
        print("case: trace, writes inside pool bricks, trace", flush=True)
        before = sim.trace(1, rays)
        ow0 = oracle.World.from_scene(s)
        assert np.array_equal(before.view(np.uint8), ow0.trace(1, rays).view(np.uint8))
        # air voxels of pool bricks, two voxels in front of primary hits
        hits = sim.trace(0, rays)
        ok = hits["face"] < 6
        p = rays["ro"][ok] + rays["rd"][ok] * (hits["t"][ok, None] - 2.0)
        v = np.floor(p).astype(np.int64)
        dx, dy, dz = s.dims
        inb = (v >= 0).all(1) & (v[:, 0] < dx) & (v[:, 1] < dy) & (v[:, 2] < dz)
        v = v[inb]
        dense = s.dense()
        air = dense[v[:, 2], v[:, 1], v[:, 0]] == 0
        cell = (v[:, 0] // 8) + s.bricks[0] * ((v[:, 1] // 8) + s.bricks[1] * (v[:, 2] // 8))
        g = s.grid[cell]
        pool_brick = (g != 0) & ((g & 0x80000000) == 0)
        v = v[air & pool_brick]
        assert v.shape[0] > 100
        key = (v[:, 0] + dx * (v[:, 1] + dy * v[:, 2])).astype(np.uint32)
        w = np.stack([key, np.full(key.shape[0], 5, np.uint32)], 1)
        sim.push_writes(w)
        sim.apply_writes()
        grid, _ = gw.read_grid()
        assert np.array_equal(kinds_of(grid), kinds_of(s.grid))  # no brick changed occupancy
        vox = sim.probe((0, 0, 0), s.dims).reshape(dz, dy, dx)
        ow1 = oracle.World.from_scene(Scene.from_dense(vox.astype(np.uint16), s.mats))
        after = sim.trace(1, rays)
        ref = ow1.trace(1, rays)
        assert np.array_equal(after.view(np.uint8), ref.view(np.uint8))
        assert not np.array_equal(after.view(np.uint8), before.view(np.uint8))
    finally:
        gw.close()


def test_debris_settles_into_voxels(floor_scene):
    """frameBegin + particleSim + writes over frames: the settled deposit
    particles become rubble voxels exactly as the oracle chain gives."""
    from pixels_amd import GpuWorld, Simulation
    from pixels_amd.scene import debris
    s = floor_scene
    gw = GpuWorld(s)
    sim = Simulation(gw)
    try:
        u = np.zeros(N.PX_UNIFORM_FLOATS, np.float32)
        u[0:3] = s.dims
        u[3], u[24], u[25] = VS, DT, G
        gw.set_uniforms(u)
        ring = 1 << 15
        sim.particles_init(ring)
        P = oracle.Particles(ring)
        recs = debris(s.dims, 20000, seed=9, speed=30.0, deposit=0.9)
        recs["pos"][:, 1] = np.minimum(recs["pos"][:, 1], 30.5)
        vox, kind = s.dense(), kinds_of(s.grid)
        total = 0
        for f in range(30):
            print("case: debris->writes frame %d" % f, flush=True)
            sim.encode(N.STAGE_FRAME_BEGIN)
            P.frame_begin()
            if f == 0:
                sim.spawn(recs)
                P.spawn(recs)
            sim.encode(N.STAGE_PARTICLES | N.STAGE_WRITES)
            w = oracle.World.from_scene(_scene_of(vox, s))
            P.step(w, DT, G, VS)
            wl = P.write_list()
            total += len(wl)
            # the pool keeps far more free bricks than these writes can take
            vox, kind, st = oracle.apply_writes(s.dims, vox, kind, wl)
            got = _state(gw, sim, s)
            _check(s, got, vox, kind)
            parts = sim.read_particles()
            assert parts.tobytes() == P.parts.tobytes()
        assert total > 1000
    finally:
        gw.close()


def _scene_of(vox, like):
    from pixels_amd.scene import Scene
    return Scene.from_dense(vox, like.mats)
