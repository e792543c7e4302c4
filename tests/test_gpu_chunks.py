"""GPU parity of island extraction and the falling chunks (px_resolve_islands,
px_chunk_step; sim.wgsl:862-923, 951-1075) with the oracle's restatement
(oracle.Chunks) chained with the debris step and the writes, frame by frame:
every decoded voxel, the brick encoding, the chunk table, the arena, the
counters, the particle ring and the impacts chunks queue.

Worlds: a 128x64x128 room (structural floor) holding a floating stone box
(a chunk that falls and lands), a fragile box over a glass pane (breaks
through, queues impacts, shatters on landing), a small blob with rubble on
top (crumbs + tumbling rubble), several boxes beyond a small max_chunks,
one component over a small component budget; and the seeded 512x256x512
scene with a floor after 64 impacts (thousands of crumbling pieces)."""
import numpy as np
import pytest

import oracle
from pixels_amd import _native as N
from tests.test_gpu_edit import kinds_of

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)
VS = np.float32(1.0 / 80.0)
G = np.float32(9.81) / VS


def room():
    dims = (128, 64, 128)
    vox = np.zeros((dims[2], dims[1], dims[0]), np.uint16)
    vox[:, :4, :] = 1                                   # structural floor
    vox[10:15, 30:35, 20:30] = 3                        # stone box, 250 voxels: a chunk
    vox[60:70, 48:52, 60:70] = 10                       # fragile box, 400 voxels
    vox[55:75, 20, 55:75] = 9                           # glass pane under it
    vox[55, 4:20, 55] = 9                               # its pillar to the floor
    vox[100:103, 30:33, 100:103] = 5                    # 27-voxel blob: crumbs
    vox[101, 33:36, 101] = 5 | 0x80                     # rubble on it
    for k in range(4):                                  # four 64-voxel boxes
        vox[20 + 10 * k:24 + 10 * k, 40:44, 100:104] = 2 + k
    vox[90:110, 10:30, 10:30] = 4                       # 8000 voxels, over the small budget below
    return dims, vox


def _uniforms(dims, frame):
    u = np.zeros(N.PX_UNIFORM_FLOATS, np.float32)
    u[0:3] = dims
    u[3] = VS
    u[24] = DT
    u[25] = G
    u[29] = frame
    return u


def _open(scene, lim):
    from pixels_amd import GpuWorld, Simulation
    gw = GpuWorld(scene, limits=lim)
    return gw, Simulation(gw)


def _state(gw, sim, dims):
    vox = sim.probe((0, 0, 0), dims).reshape(dims[2], dims[1], dims[0])
    grid, _ = gw.read_grid()
    ctrl, args, _ = sim.read_ctrl()
    return vox, grid, ctrl, args


def _read_chunks(gw, max_chunks):
    import ctypes as C
    table = np.zeros(max_chunks * 16, np.uint32)
    arena = np.zeros(2 * (1 << 20), np.uint32)
    n = C.c_uint32(0)
    gw.check(gw.lib.px_read_chunks(gw.ctx, N.ptr(table), N.ptr(arena), 1 << 20, C.byref(n)))
    return table, arena[:2 * n.value].copy()


def _compare(gw, sim, dims, ch, vox, kind, P, what, max_chunks):
    gvox, grid, ctrl, args = _state(gw, sim, dims)
    assert np.array_equal(gvox, vox.astype(np.uint32)), (what, "voxels")
    assert np.array_equal(kinds_of(grid), kind), (what, "bricks")
    table, arena = _read_chunks(gw, max_chunks)
    n = min(int(ch.ctr[oracle.R_C_CHUNKS]), max_chunks)
    assert np.array_equal(table[:16 * n], ch.chunks[:16 * n]), (what, "chunk table")
    na = min(int(ch.ctr[oracle.R_C_ARENA]), ch.max_arena)
    assert np.array_equal(arena, ch.arena[:2 * na]), (what, "arena")
    c = N.CTRL
    assert int(ctrl[c["C_CHUNKS"]]) == int(ch.ctr[oracle.R_C_CHUNKS]), what
    assert int(ctrl[c["C_ARENA"]]) == int(ch.ctr[oracle.R_C_ARENA]), what
    assert int(ctrl[c["S_ISLANDS"]]) == int(ch.ctr[oracle.R_S_ISLANDS]), what
    assert int(ctrl[c["S_BUDGET"]]) == int(ch.ctr[oracle.R_S_BUDGET]), what
    parts = sim.read_particles()
    assert parts.tobytes() == P.parts.tobytes(), (what, "particles")
    assert (int(ctrl[c["C_SPAWNED"]]), int(ctrl[c["C_PCURSOR"]])) == (int(P.ctr[oracle.P_SPAWNED]),
                                                                      int(P.ctr[oracle.P_PCURSOR])), what
    return ctrl


@pytest.mark.parametrize("max_chunks,budget", [(256, 400000), (3, 5000)])
def test_room_extraction_and_falls(max_chunks, budget):
    from pixels_amd.scene import Scene
    dims, vox0 = room()
    scene = Scene.from_dense(vox0)
    lim = N.default_limits()
    lim.max_chunks = max_chunks
    lim.component_budget = budget
    gw, sim = _open(scene, lim)
    try:
        sim.particles_init(1 << 14)
        P = oracle.Particles(1 << 14, max_writes=int(lim.max_writes))
        ch = oracle.Chunks(dims, scene.mats, max_chunks=max_chunks, component_budget=budget)
        gw.set_uniforms(_uniforms(dims, 1))
        print("case: room max_chunks %d budget %d: island pass + resolve" % (max_chunks, budget), flush=True)
        sim.encode(N.STAGE_FRAME_BEGIN)
        P.frame_begin()
        sim.island_pass()
        sim.encode(N.STAGE_RESOLVE)
        keys, cnt, nl = oracle.islands(dims, vox0, scene.mats)
        vox, kind, sp = ch.resolve(vox0, kinds_of(scene.grid), keys, cnt, VS, 1.0)
        P.spawn(sp)
        _compare(gw, sim, dims, ch, vox, kind, P, "resolve", max_chunks)
        assert int(ch.ctr[oracle.R_C_ARENA]) > 0 and len(sp) > 0
        landed = impacts = 0
        for f in range(2, 90):
            gw.set_uniforms(_uniforms(dims, f))
            print("case: room frame %d" % f, flush=True)
            sim.encode(N.STAGE_FRAME_BEGIN | N.STAGE_CHUNKS | N.STAGE_PARTICLES | N.STAGE_WRITES)
            P.frame_begin()
            ch.ctr[oracle.R_C_WRITES] = 0
            sp = ch.frame(vox, kind, DT, G, VS, f)
            P.spawn(sp)
            cw = ch.write_list()
            P.writes[:cw.size] = cw.ravel()
            P.ctr[oracle.P_WRITES] = ch.ctr[oracle.R_C_WRITES]
            P.step(oracle.World.from_scene(Scene.from_dense(vox, scene.mats)), DT, G, VS)
            wl = P.write_list()
            landed += len(cw)
            vox, kind, _ = oracle.apply_writes(dims, vox, kind, wl)
            _compare(gw, sim, dims, ch, vox, kind, P, "frame %d" % f, max_chunks)
            out = np.zeros(64 * 16, np.float32)
            import ctypes as C
            n = C.c_uint32(0)
            gw.check(gw.lib.px_read_device_impacts(gw.ctx, N.ptr(out, C.c_float), 64, C.byref(n)))
            assert n.value == int(ch.ctr[oracle.R_C_IMPACTS]), f
            m = min(n.value, 64)
            assert np.array_equal(out[:16 * m], ch.impact_list().ravel()[:16 * m]), f
            impacts = max(impacts, m)
        assert landed > 0
        if max_chunks == 256:
            assert impacts > 0
    finally:
        gw.close()


def test_floor_scene_resolve(floor_scene):
    """After 64 impacts on the seeded scene: every unsupported piece leaves
    the world (chunks into the arena, crumbs and rubble into the ring)."""
    from pixels_amd.world import impacts_from_hits
    s = floor_scene
    lim = N.default_limits()
    gw, sim = _open(s, lim)
    try:
        rng = np.random.default_rng(8)
        rays = s.camera_rays(640, 360, 90.0, seed=3)
        pick = rays[rng.permutation(rays.shape[0])[:400]]
        hits = sim.trace(0, pick)
        ok = hits["face"] != 0xFF
        pick, hits = pick[ok][:64], hits[ok][:64]
        r = rng.uniform(8, 96, pick.shape[0]).astype(np.float32)
        impacts = impacts_from_hits(s.dims, pick, hits, r)
        sim.particles_init(1 << 18)
        P = oracle.Particles(1 << 18)
        gw.set_uniforms(_uniforms(s.dims, 5))
        print("case: floor scene island pass + resolve", flush=True)
        sim.push_impacts(impacts)
        sim.encode(N.STAGE_FRAME_BEGIN | N.STAGE_EDIT | N.STAGE_RESOLVE)
        P.frame_begin()
        r = oracle.edit_pass(s.dims, s.dense(), kinds_of(s.grid), s.mats, impacts, vs=VS, frame=5.0)
        P.spawn(r.spawns)  # carve and shard debris, rubble drops
        stats = r.stats
        ch = oracle.Chunks(s.dims, s.mats)
        vox, kind, sp = ch.resolve(r.vox, r.kind, r.keys, r.counts, VS, 5.0)
        P.spawn(sp)
        ctrl = _compare(gw, sim, s.dims, ch, vox, kind, P, "floor resolve", 256)
        c = N.CTRL
        assert int(ctrl[c["S_DESTROYED"]]) == int(stats[0]) + int(ch.ctr[oracle.R_S_DESTROYED])
        assert int(ctrl[c["S_BILL"]]) == int(stats[1]) + int(ch.ctr[oracle.R_S_BILL])
        assert len(sp) > 1000
        print("floor scene: %d labels, %d chunks, %d arena voxels, %d + %d debris records" %
              (r.n, int(ch.ctr[oracle.R_S_ISLANDS]), int(ch.ctr[oracle.R_C_ARENA]), len(r.spawns), len(sp)),
              flush=True)
    finally:
        gw.close()


def test_resolve_without_island_pass_fails(small_scene):
    from pixels_amd import PxError
    lim = N.default_limits()
    gw, sim = _open(small_scene, lim)
    try:
        with pytest.raises(PxError) as e:
            gw.check(gw.lib.px_resolve_islands(gw.ctx))
        assert e.value.code == -4
        sim.encode(N.STAGE_RESOLVE)  # the stage bit skips when there is nothing to resolve
    finally:
        gw.close()
