"""GPU parity of the edit pass's side effects and of the seeded search
against the oracle (o_edit_pass), bit for bit:

* carveEffects / shardEffects (sim.wgsl:261-284, 485-524): rubble stacked on
  carved and flooded voxels tumbles, and the carve and shard debris is
  spawned, in the fixed order (carved voxels, then flooded ones, each in
  cell / local-index order): the whole particle ring, C_SPAWNED, C_PCURSOR,
  the bill and destroyed counters and every voxel are compared;
* the seeded search (PX_ISLANDS_SEEDED, sim.wgsl:620-882): labels equal the
  oracle's seeded labels on a room with a floating block no damage touches
  (it stays), on the 512x256x512 scene with its floor after 64 impacts, with
  the node budget tripped, and through the resolve, where a tiny component
  budget trips S_BUDGET;
* the pass is recorded whatever the device holds: with nothing queued it
  leaves the world and the last labels alone; device-side impacts (falling
  chunks) and pending landing seeds start it with no host round trip.
"""
import numpy as np
import pytest

import oracle
from pixels_amd import _native as N
from tests.test_gpu_edit import kinds_of

pytestmark = pytest.mark.gpu

VS = np.float32(1.0 / 80.0)


def _uniforms(dims, frame):
    u = np.zeros(N.PX_UNIFORM_FLOATS, np.float32)
    u[0:3] = dims
    u[3] = VS
    u[24] = np.float32(1.0 / 60.0)
    u[25] = np.float32(9.81) / VS
    u[29] = frame
    return u


def _impact(c, r, fr, impulse=0.0, n=(0, 0, 0), d=(0, 0, 0)):
    im = np.zeros(16, np.float32)
    im[0:3], im[3], im[4], im[5] = c, r, fr, impulse
    im[6:9], im[9:12] = n, d
    return im


def _effects_world(seed):
    rng = np.random.default_rng(seed)
    dims = (128, 64, 96)
    vox = np.zeros(dims[::-1], np.uint16)
    vox[:, 0:4, :] = 1
    for _ in range(14):  # blocks of several materials with rubble piles on top
        z, y, x = rng.integers([4, 4, 4], [80, 36, 112])
        h = rng.integers(3, 10, 3)
        vox[z:z + h[0], y:y + h[1], x:x + h[2]] = rng.choice([2, 3, 5, 8, 10, 14, 16])
        vox[z:z + h[0], y + h[1]:y + h[1] + rng.integers(1, 16), x:x + h[2]:2] = rng.choice([3, 5, 14]) | 0x80
    for _ in range(6):  # windows and shelves
        z0, y0, x0 = rng.integers([4, 6, 4], [70, 30, 100])
        vox[z0:z0 + 14, y0:y0 + 16, x0:x0 + 2] = 9
        vox[z0 + 3:z0 + 10, y0 + 5, x0 + 2:x0 + 14] = 12
        vox[z0 + 4:z0 + 8, y0 + 16:y0 + 20, x0:x0 + 2] = 2 | 0x80  # rubble on the window
    solid = np.argwhere((vox != 0) & (vox != 1))
    pick = solid[rng.permutation(solid.shape[0])[:24]]
    imps = [_impact((p[2] + 0.5, p[1] + 0.5, p[0] + 0.5), rng.uniform(2, 9), rng.uniform(2, 12),
                    impulse=rng.uniform(50, 400), n=rng.normal(size=3) * 0.5, d=rng.normal(size=3) * 0.5)
            for p in pick]
    return dims, vox, np.array(imps, np.float32)


def _run(scene, lim, imps, frame, ring=1 << 16, setup=None):
    from pixels_amd import GpuWorld, Simulation
    gw = GpuWorld(scene, limits=lim)
    sim = Simulation(gw)
    gw.set_uniforms(_uniforms(scene.dims, frame))
    sim.particles_init(ring)
    if setup:
        setup(gw, sim)
    sim.encode(N.STAGE_FRAME_BEGIN)
    if len(imps):
        sim.push_impacts(imps)
    sim.island_pass()
    return gw, sim


def _compare_pass(gw, sim, scene, r, P, what):
    dims = scene.dims
    vox = sim.probe((0, 0, 0), dims).reshape(dims[2], dims[1], dims[0])
    assert np.array_equal(vox, r.vox.astype(np.uint32)), (what, "voxels")
    grid, _ = gw.read_grid()
    assert np.array_equal(kinds_of(grid), r.kind), (what, "bricks")
    keys, cnt, n = sim.island_labels()
    assert n == r.n and np.array_equal(keys, r.keys) and np.array_equal(cnt, r.counts), (what, "labels")
    parts = sim.read_particles()
    if parts.tobytes() != P.parts.tobytes():
        a = parts.view(np.uint32).reshape(len(parts), -1)
        b = P.parts.view(np.uint32).reshape(len(P.parts), -1)
        bad = np.nonzero((a != b).any(1))[0]
        ctrl, _, _ = sim.read_ctrl()
        print("%s: %d of %d ring slots differ, first %s; gpu C_SPAWNED %d C_PCURSOR %d, oracle %d %d, calls %d" %
              (what, bad.size, len(parts), bad[:8], ctrl[N.CTRL["C_SPAWNED"]], ctrl[N.CTRL["C_PCURSOR"]],
               P.ctr[oracle.P_SPAWNED], P.ctr[oracle.P_PCURSOR], r.spawn_calls), flush=True)
        for k in bad[:4]:
            print(" slot", k, "gpu", parts[k], "oracle", P.parts[k], flush=True)
        assert False, (what, "particle ring")
    ctrl, _, _ = sim.read_ctrl()
    c = N.CTRL
    assert (int(ctrl[c["C_SPAWNED"]]), int(ctrl[c["C_PCURSOR"]])) == (int(P.ctr[oracle.P_SPAWNED]),
                                                                      int(P.ctr[oracle.P_PCURSOR])), what
    assert int(ctrl[c["S_DESTROYED"]]) == int(r.stats[0]), what
    assert int(ctrl[c["S_BILL"]]) == int(r.stats[1]), what
    return ctrl


@pytest.mark.parametrize("seed,ring", [(1, 1 << 16), (2, 1 << 12)])
def test_carve_and_shard_effects_bit_exact(seed, ring):
    from pixels_amd import Scene
    dims, vox, imps = _effects_world(seed)
    s = Scene.from_dense(vox)
    lim = N.default_limits()
    print("case: effects world seed %d ring %d" % (seed, ring), flush=True)
    gw, sim = _run(s, lim, imps, 11.0, ring=ring)
    try:
        r = oracle.edit_pass(dims, vox, kinds_of(s.grid), s.mats, imps, vs=VS, frame=11.0)
        P = oracle.Particles(ring, spawn_budget=int(lim.spawn_budget))
        P.frame_begin()
        P.spawn(r.spawns)
        _compare_pass(gw, sim, s, r, P, "effects")
        assert int(r.stats[7]) > 20 and int(r.stats[4]) > 0  # rubble dropped, panes flooded
        assert r.spawn_calls > 100
        print("effects: %d records, %d rubble, %d flooded" % (r.spawn_calls, r.stats[7], r.stats[4]), flush=True)
    finally:
        gw.close()


def test_effects_on_the_floor_scene_over_the_spawn_budget(floor_scene):
    """The 64-impact pass on the 512x256x512 scene spawns more records than a
    20,000-record frame budget: the first 20,000 in the fixed order are
    stored."""
    from pixels_amd.world import impacts_from_hits
    from pixels_amd import GpuWorld, Simulation
    s = floor_scene
    gw0 = GpuWorld(s)
    sim0 = Simulation(gw0)
    rng = np.random.default_rng(8)
    rays = s.camera_rays(640, 360, 90.0, seed=3)
    pick = rays[rng.permutation(rays.shape[0])[:400]]
    hits = sim0.trace(0, pick)
    gw0.close()
    ok = hits["face"] != 0xFF
    imps = impacts_from_hits(s.dims, pick[ok][:64], hits[ok][:64], rng.uniform(8, 96, 64).astype(np.float32),
                             impulse=200.0)
    lim = N.default_limits()
    lim.spawn_budget = 20000
    print("case: floor scene effects, 64 impacts", flush=True)
    gw, sim = _run(s, lim, imps, 4.0, ring=1 << 18)
    try:
        r = oracle.edit_pass(s.dims, s.dense(), kinds_of(s.grid), s.mats, imps, vs=VS, frame=4.0)
        P = oracle.Particles(1 << 18, spawn_budget=int(lim.spawn_budget))
        P.frame_begin()
        P.spawn(r.spawns)
        ctrl = _compare_pass(gw, sim, s, r, P, "floor scene")
        assert r.spawn_calls > int(lim.spawn_budget)
        assert int(ctrl[N.CTRL["C_PCURSOR"]]) == int(lim.spawn_budget)
    finally:
        gw.close()


def _seeded_room():
    dims = (96, 64, 96)
    vox = np.zeros(dims[::-1], np.uint16)
    vox[:, 0:4, :] = 1
    vox[40:44, 4:40, 40:44] = 2              # a pillar carrying a table top
    vox[30:54, 40:44, 30:54] = 3
    vox[60:70, 30:40, 60:70] = 5             # a floating block no impact touches
    vox[10:14, 4:30, 10:14] = 8              # a second pillar with its own top
    vox[8:16, 30:34, 8:16] = 13
    vox[70:80, 20:22, 10:30] = 9             # a pane on a post
    vox[74:76, 4:20, 10:12] = 4
    vox[41, 44:50, 41] = 3 | 0x80            # rubble on the table
    imps = np.array([_impact((42.0, 20.0, 42.0), 6.0, 6.0, 100.0), _impact((12.0, 15.0, 12.0), 8.0, 8.0, 100.0),
                     _impact((11.0, 12.0, 75.0), 6.0, 6.0, 100.0)], np.float32)
    return dims, vox, imps


def _seeded_limits(**kw):
    lim = N.default_limits()
    lim.island_mode = N.ISLANDS_SEEDED
    for k, v in kw.items():
        setattr(lim, k, v)
    return lim


def test_seeded_room_leaves_the_untouched_block():
    from pixels_amd import Scene
    dims, vox, imps = _seeded_room()
    s = Scene.from_dense(vox)
    lim = _seeded_limits()
    print("case: seeded room", flush=True)
    gw, sim = _run(s, lim, imps, 3.0)
    try:
        r = oracle.edit_pass(dims, vox, kinds_of(s.grid), s.mats, imps, mode=1, vs=VS, frame=3.0,
                             node_budget=(int(lim.hash_size) * 7) // 10)
        P = oracle.Particles(1 << 16, spawn_budget=int(lim.spawn_budget))
        P.frame_begin()
        P.spawn(r.spawns)
        ctrl = _compare_pass(gw, sim, s, r, P, "seeded room")
        block = 60 + 96 * (30 + 64 * 60)
        assert r.n >= 2 and block not in set(r.keys.tolist())
        assert int(ctrl[N.CTRL["C_SEED_N"]]) == int(r.stats[5]) > 0
        # the global form labels the block too
        r_all = oracle.edit_pass(dims, vox, kinds_of(s.grid), s.mats, imps, mode=0)
        assert block in set(r_all.keys.tolist())
        # extraction leaves the block in place
        sim.encode(N.STAGE_RESOLVE)
        after = sim.probe((60, 30, 60), (10, 10, 10))
        assert np.all(after == 5)
    finally:
        gw.close()


@pytest.mark.parametrize("budget_frac", [None, 0.5])
def test_seeded_floor_scene_and_node_budget(floor_scene, budget_frac):
    from pixels_amd import GpuWorld, Simulation
    from pixels_amd.world import impacts_from_hits
    s = floor_scene
    gw0 = GpuWorld(s)
    sim0 = Simulation(gw0)
    rng = np.random.default_rng(8)
    rays = s.camera_rays(640, 360, 90.0, seed=3)
    pick = rays[rng.permutation(rays.shape[0])[:400]]
    hits = sim0.trace(0, pick)
    gw0.close()
    ok = hits["face"] != 0xFF
    imps = impacts_from_hits(s.dims, pick[ok][:64], hits[ok][:64], rng.uniform(8, 96, 64).astype(np.float32))
    kind = kinds_of(s.grid)
    full = oracle.edit_pass(s.dims, s.dense(), kind, s.mats, imps, mode=1, vs=VS, frame=2.0)
    total = int(full.counts.astype(np.int64).sum())
    if budget_frac is None:
        hash_size = 1 << 26  # 7/10 of it holds every floating voxel
    else:
        hash_size = int(total * budget_frac / 0.7)
    budget = (hash_size * 7) // 10
    lim = _seeded_limits(hash_size=hash_size)
    print("case: seeded floor scene, hash_size %d (node budget %d of %d floating voxels)" % (hash_size, budget, total),
          flush=True)
    gw, sim = _run(s, lim, imps, 2.0, ring=1 << 18)
    try:
        r = oracle.edit_pass(s.dims, s.dense(), kind, s.mats, imps, mode=1, vs=VS, frame=2.0, node_budget=budget)
        P = oracle.Particles(1 << 18, spawn_budget=int(lim.spawn_budget))
        P.frame_begin()
        P.spawn(r.spawns)
        _compare_pass(gw, sim, s, r, P, "seeded floor scene")
        g_all = oracle.edit_pass(s.dims, s.dense(), kind, s.mats, imps, mode=0)
        assert 0 < r.n <= g_all.n
        if budget_frac is None:
            assert int(r.stats[6]) == 0 and np.array_equal(r.keys, full.keys)
        else:
            assert int(r.stats[6]) > 0 and r.n < full.n
            assert np.array_equal(r.keys, full.keys[:r.n])
    finally:
        gw.close()


def test_seeded_resolve_trips_component_budget():
    """Seeded labels through the extraction: a component over the component
    budget stays in the world and counts in S_BUDGET."""
    from pixels_amd import Scene
    dims, vox, imps = _seeded_room()
    s = Scene.from_dense(vox)
    lim = _seeded_limits(component_budget=600)
    gw, sim = _run(s, lim, imps, 3.0)
    try:
        keys, cnt, n = sim.island_labels()
        big = int((cnt > 600).sum())
        assert big >= 1
        sim.encode(N.STAGE_RESOLVE)
        ctrl, _, _ = sim.read_ctrl()
        assert int(ctrl[N.CTRL["S_BUDGET"]]) == big
        r = oracle.edit_pass(dims, vox, kinds_of(s.grid), s.mats, imps, mode=1, vs=VS, frame=3.0,
                             node_budget=(int(lim.hash_size) * 7) // 10)
        ch = oracle.Chunks(dims, s.mats, component_budget=600)
        v2, k2, _ = ch.resolve(r.vox, r.kind, r.keys, r.counts, VS, 3.0)
        got = sim.probe((0, 0, 0), dims).reshape(dims[2], dims[1], dims[0])
        assert np.array_equal(got, v2.astype(np.uint32))
        assert int(ch.ctr[oracle.R_S_BUDGET]) == big
    finally:
        gw.close()


def test_pass_without_work_keeps_world_and_labels(floor_scene):
    """px_sim_step(PX_STAGE_EDIT) with nothing queued runs nothing on the
    device: the world and the labels of the last pass stay."""
    from pixels_amd import GpuWorld, Simulation
    s = floor_scene
    gw = GpuWorld(s)
    try:
        sim = Simulation(gw)
        sim.island_pass()  # labels of the world as built
        k0, c0, n0 = sim.island_labels()
        assert n0 > 0
        for _ in range(3):
            sim.encode(N.STAGE_EDIT)
        k1, c1, n1 = sim.island_labels()
        assert n1 == n0 and np.array_equal(k1, k0) and np.array_equal(c1, c0)
        grid, _ = gw.read_grid()
        assert np.array_equal(grid, s.grid)
        ctrl, _, _ = sim.read_ctrl()
        assert int(ctrl[N.CTRL["S_SEARCHES"]]) == 1
    finally:
        gw.close()


def test_device_impacts_and_landing_seeds_start_the_pass():
    """A box knocked loose falls 10 voxels onto a pedestal with a fragile top:
    it breaks through (impacts queued on the device), then lands on the
    fragile layer (landing seeds, pushSeed :1068-1072).  The next
    px_sim_step(EDIT) takes both without a host read."""
    from pixels_amd import Scene
    dims = (96, 64, 96)
    vox = np.zeros(dims[::-1], np.uint16)
    vox[:, 0:4, :] = 1
    vox[40:60, 4:19, 40:60] = 2      # pedestal
    vox[40:60, 19:21, 40:60] = 10    # its fragile top layer
    vox[45:55, 31:37, 45:55] = 3     # the box, held by nothing
    s = Scene.from_dense(vox)
    lim = _seeded_limits()
    imps = np.array([_impact((50.0, 36.9, 50.0), 1.5, 1.5)], np.float32)  # a chip off the box's top: seeds it
    gw, sim = _run(s, lim, imps, 1.0)
    try:
        keys, cnt, n = sim.island_labels()
        assert n == 1 and int(keys[0]) == 45 + 96 * (31 + 64 * 45)
        sim.encode(N.STAGE_RESOLVE)
        queued = pend = 0
        for f in range(2, 60):
            gw.set_uniforms(_uniforms(dims, f))
            sim.encode(N.STAGE_FRAME_BEGIN | N.STAGE_CHUNKS | N.STAGE_PARTICLES | N.STAGE_WRITES)
            ctrl, _, _ = sim.read_ctrl()
            queued = max(queued, int(ctrl[N.CTRL["C_IMPACTS"]]))
            pend = max(pend, int(ctrl[N.CTRL["C_PENDING"]]))
        print("case: device impacts %d, pending seeds %d" % (queued, pend), flush=True)
        assert queued > 0 and pend > 0
        ctrl, _, _ = sim.read_ctrl()
        searches0 = int(ctrl[N.CTRL["S_SEARCHES"]])
        sim.encode(N.STAGE_EDIT)
        ctrl, _, _ = sim.read_ctrl()
        assert int(ctrl[N.CTRL["S_SEARCHES"]]) == searches0 + 1
        assert int(ctrl[N.CTRL["C_IMPACTS"]]) == 0 and int(ctrl[N.CTRL["C_PENDING"]]) == 0
        sim.run_to_quiescence()
    finally:
        gw.close()
