"""GPU parity of the shoot stage (PX_STAGE_SHOOT, sim.wgsl:371-403) against
the oracle (o_shoot, then o_edit_pass with the shot's flood entries), bit for
bit:

* after FRAME_BEGIN | SHOOT: the impact queue (C_IMPACTS and every record),
  the flash ring and C_FLASH, every voxel (struck panes and the rubble above
  them taken), the bricks, the particle ring with C_SPAWNED / C_PCURSOR, the
  bill and destroyed counters;
* after the EDIT stage that follows: every voxel and brick, the labels, the
  particle ring and the counters, in both island modes;
* cases: a spread volley at a window with rubble on it, a straight shot at
  the pane's top row, more than 16 pellets (16 fire), no pellets (nothing
  happens and the edit stage leaves the world alone), and the whole frame in
  one px_sim_step call.
"""
import numpy as np
import pytest

import oracle
from pixels_amd import _native as N
from tests.test_gpu_edit import kinds_of

pytestmark = pytest.mark.gpu

VS = np.float32(1.0 / 80.0)
DIMS = (128, 64, 96)


def _world():
    rng = np.random.default_rng(4)
    vox = np.zeros(DIMS[::-1], np.uint16)  # z, y, x
    vox[:, 0:4, :] = 1
    for _ in range(10):
        z, y, x = rng.integers([4, 4, 4], [80, 30, 112])
        h = rng.integers(3, 10, 3)
        vox[z:z + h[0], y:y + h[1], x:x + h[2]] = rng.choice([2, 3, 5, 8, 10, 14, 16])
    vox[20:70, 4:41, 60:62] = 9            # a window, top row y = 40
    vox[44, 41:47, 60] = 2 | 0x80         # rubble on the struck column
    vox[40:50, 41:44, 61] = 3 | 0x80
    vox[30:40, 20, 62:80] = 12            # a glass shelf joined to the window's side
    return vox


def _uniforms(pellets, r, fr, impulse, spread=0.0, seed=0.0, frame=9.0):
    u = np.zeros(N.PX_UNIFORM_FLOATS, np.float32)
    u[0:3] = DIMS
    u[3] = VS
    u[24] = np.float32(1.0 / 60.0)
    u[25] = np.float32(9.81) / VS
    u[28], u[29] = 5.5, frame
    u[40:44] = (pellets, r, fr, impulse)
    u[44:46] = (spread, seed)
    return u


def _camera(eye_vox, look):
    cam = np.zeros(N.PX_CAMERA_FLOATS, np.float32)
    cam[32:35] = np.asarray(eye_vox, np.float32) * VS
    d = np.asarray(look, np.float64)
    cam[36:39] = d / np.linalg.norm(d)
    return cam


CASES = {
    "volley": (_uniforms(10, 3.0, 6.0, 300.0, spread=0.08, seed=77.0), _camera((20.5, 30.5, 45.5), (1, 0.05, 0))),
    "top_row": (_uniforms(3, 4.0, 2.0, 150.0), _camera((30.5, 40.5, 44.5), (1, 0, 0))),
    "over_16": (_uniforms(40, 2.0, 3.0, 200.0, spread=0.5, seed=3.0), _camera((40.5, 25.5, 40.5), (1, -0.2, 0.3))),
    "none": (_uniforms(0, 3.0, 6.0, 300.0), _camera((20.5, 30.5, 45.5), (1, 0, 0))),
}


def _gpu(scene, u, cam, mode, ring):
    from pixels_amd import GpuWorld, Simulation
    lim = N.default_limits()
    lim.island_mode = mode
    gw = GpuWorld(scene, limits=lim)
    sim = Simulation(gw)
    gw.set_uniforms(u)
    sim.particles_init(ring)
    sim.set_camera(cam)
    return gw, sim, lim


def _check_ring(sim, P, what):
    parts = sim.read_particles()
    assert parts.tobytes() == P.parts.tobytes(), (what, "particle ring")
    ctrl, _, _ = sim.read_ctrl()
    c = N.CTRL
    assert int(ctrl[c["C_SPAWNED"]]) == int(P.ctr[oracle.P_SPAWNED]), what
    assert int(ctrl[c["C_PCURSOR"]]) == int(P.ctr[oracle.P_PCURSOR]), what
    return ctrl


@pytest.mark.parametrize("mode", [N.ISLANDS_ALL, N.ISLANDS_SEEDED])
@pytest.mark.parametrize("case", sorted(CASES))
def test_shoot_then_edit_bit_exact(case, mode):
    from pixels_amd import Scene
    vox = _world()
    s = Scene.from_dense(vox)
    u, cam = CASES[case]
    ring = 1 << 14
    print("case: shoot %s mode %d" % (case, mode), flush=True)
    gw, sim, lim = _gpu(s, u, cam, mode, ring)
    try:
        sim.encode(N.STAGE_FRAME_BEGIN | N.STAGE_SHOOT)
        imps, n = sim.read_device_impacts()
        fl = sim.read_flashes()
        r = oracle.shoot(DIMS, vox, kinds_of(s.grid), s.mats, u, cam)
        assert n == r.n_impacts, (case, "C_IMPACTS")
        assert imps[:, :12].tobytes() == r.impacts[:min(n, 64), :12].tobytes(), (case, "impact records")
        assert fl.tobytes() == r.flashes.tobytes(), (case, "flashes")
        ctrl, _, _ = sim.read_ctrl()
        c = N.CTRL
        assert int(ctrl[c["C_FLASH"]]) == r.flash_ctr
        assert int(ctrl[c["S_DESTROYED"]]) == int(r.stats[0]) and int(ctrl[c["S_BILL"]]) == int(r.stats[1])
        got = sim.probe((0, 0, 0), DIMS).reshape(vox.shape)
        assert np.array_equal(got, r.vox.astype(np.uint32)), (case, "voxels after shoot")
        grid, _ = gw.read_grid()
        assert np.array_equal(kinds_of(grid), r.kind)
        P = oracle.Particles(ring, spawn_budget=int(lim.spawn_budget))
        P.frame_begin()
        P.spawn(r.spawns)
        _check_ring(sim, P, case + " shoot")
        if case == "volley":
            assert r.n_flood >= 1 and r.stats[2] >= 8
        if case == "top_row":
            assert r.n_flood == 1 and r.spawn_calls >= 6  # the rubble column tumbles
        if case == "over_16":
            assert r.stats[2] <= 16 and r.n_impacts <= 16 and r.n_impacts >= 8
        if case == "none":
            assert n == 0 and r.flash_ctr == 0
        # the edit stage consumes the queue, the flood entries and the dirty cells
        sim.encode(N.STAGE_EDIT)
        e = oracle.edit_pass(DIMS, r.vox, r.kind, s.mats, r.impacts[:min(n, 64)], mode=mode, vs=VS,
                             frame=float(u[29]), flood=r.flood, dirty=r.dirty)
        got = sim.probe((0, 0, 0), DIMS).reshape(vox.shape)
        assert np.array_equal(got, e.vox.astype(np.uint32)), (case, "voxels after edit")
        grid, _ = gw.read_grid()
        assert np.array_equal(kinds_of(grid), e.kind), (case, "bricks after edit")
        if n:
            keys, cnt, nl = sim.island_labels()
            assert nl == e.n and np.array_equal(keys, e.keys) and np.array_equal(cnt, e.counts), (case, "labels")
            if case in ("volley", "top_row"):
                assert not (e.vox == 9).any()  # the window and the shelf of another id
                assert (e.vox == 12).any()
        P.spawn(e.spawns)
        ctrl = _check_ring(sim, P, case + " edit")
        assert int(ctrl[c["S_DESTROYED"]]) == int(r.stats[0] + e.stats[0])
        assert int(ctrl[c["C_IMPACTS"]]) == 0
    finally:
        gw.close()


def test_one_frame_call_equals_the_stages():
    """FRAME_BEGIN | SHOOT | EDIT in one px_sim_step call gives the same world,
    ring and counters as the stages called one by one."""
    from pixels_amd import Scene
    vox = _world()
    s = Scene.from_dense(vox)
    u, cam = CASES["volley"]
    outs = []
    for split in (False, True):
        print("case: one frame call, split=%s" % split, flush=True)
        gw, sim, lim = _gpu(s, u, cam, N.ISLANDS_ALL, 1 << 14)
        try:
            if split:
                sim.encode(N.STAGE_FRAME_BEGIN)
                sim.encode(N.STAGE_SHOOT)
                sim.encode(N.STAGE_EDIT)
            else:
                sim.encode(N.STAGE_FRAME_BEGIN | N.STAGE_SHOOT | N.STAGE_EDIT)
            ctrl, _, _ = sim.read_ctrl()
            outs.append((sim.probe((0, 0, 0), DIMS).tobytes(), sim.read_particles().tobytes(), ctrl.tobytes(),
                         sim.read_flashes().tobytes()))
        finally:
            gw.close()
    assert outs[0] == outs[1]
