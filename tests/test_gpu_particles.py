"""GPU parity of the debris step (px_particles.hip) with the oracle's serial
restatement (oracle.Particles, sim.wgsl:166-170, 363-369, 1088-1173):
after every frame of frameBegin + spawn + particleSim the ring (all 32 bytes
of every slot), the alive list, the write list and the counters C_SPAWNED,
C_PCURSOR, C_WRITES and args[A_DRAWP + 1] are bit-identical.  Cases: the
reference limits, a small spawn budget with a write list that overflows, the
smallest ring (2^12) wrapping inside one spawn call, the largest ring (2^21),
the stage mask of px_sim_step, and the state / argument errors."""
import numpy as np
import pytest

import oracle
from pixels_amd import _native as N
from pixels_amd.scene import debris

pytestmark = pytest.mark.gpu

DT = np.float32(1.0 / 60.0)
VS = np.float32(1.0 / 80.0)
G = np.float32(9.81 / VS)
A_DRAWP = 36


def _uniforms(dims, dt=DT):
    u = np.zeros(N.PX_UNIFORM_FLOATS, np.float32)
    u[0:3] = dims
    u[3] = VS
    u[24] = dt
    u[25] = G
    return u


def _open(scene, budget=60000, max_writes=131072):
    from pixels_amd import GpuWorld, Simulation
    lim = N.default_limits()
    lim.spawn_budget = budget
    lim.max_writes = max_writes
    gw = GpuWorld(scene, limits=lim)
    gw.set_uniforms(_uniforms(scene.dims))
    return gw, Simulation(gw)


def _compare(sim, P, what):
    got = sim.read_particles()
    assert got.tobytes() == P.parts.tobytes(), (what, _first_diff(got, P.parts))
    assert np.array_equal(sim.read_alive(), P.alive_list()), what
    assert np.array_equal(sim.read_writes(), P.write_list()), what
    ctrl, args, _ = sim.read_ctrl()
    c = N.CTRL
    assert (ctrl[c["C_SPAWNED"]], ctrl[c["C_PCURSOR"]], ctrl[c["C_WRITES"]], args[A_DRAWP + 1]) == \
        tuple(int(x) for x in P.ctr), what


def _first_diff(a, b):
    ab, bb = a.view(np.uint32).reshape(len(a), 8), b.view(np.uint32).reshape(len(b), 8)
    rows = np.nonzero(np.any(ab != bb, axis=1))[0]
    if not len(rows):
        return None
    r = rows[0]
    return dict(slot=int(r), n_diff=int(len(rows)), gpu=a[r], oracle=b[r])


@pytest.mark.parametrize("budget,max_writes,ring", [(60000, 131072, 1 << 14), (2000, 40, 1 << 14)])
def test_frames_bit_exact(small_scene, budget, max_writes, ring):
    s = small_scene
    w = oracle.World.from_scene(s)
    gw, sim = _open(s, budget, max_writes)
    try:
        sim.particles_init(ring)
        P = oracle.Particles(ring, spawn_budget=budget, max_writes=max_writes)
        for f in range(24):
            print("case: debris frame %d budget %d max_writes %d" % (f, budget, max_writes), flush=True)
            recs = debris(s.dims, 3000, seed=100 + f, speed=120.0)
            sim.frame_begin()
            P.frame_begin()
            sim.spawn(recs)
            P.spawn(recs)
            sim.particle_step()
            P.step(w, DT, G, VS)
            _compare(sim, P, "frame %d" % f)
        assert int(P.ctr[oracle.P_ALIVE]) > 1000
    finally:
        gw.close()


def test_settling_writes_and_overflow(floor_scene):
    """Debris over the structural floor: most particles settle within the
    frames, the write list overflows max_writes on the busy frames."""
    s = floor_scene
    w = oracle.World.from_scene(s)
    gw, sim = _open(s, 60000, 500)
    try:
        ring = 1 << 15
        sim.particles_init(ring)
        P = oracle.Particles(ring, max_writes=500)
        recs = debris(s.dims, 30000, seed=7, speed=40.0)
        recs["pos"][:, 1] = np.minimum(recs["pos"][:, 1], 40.5)
        overflowed = 0
        for f in range(40):
            print("case: floor frame %d" % f, flush=True)
            sim.frame_begin()
            P.frame_begin()
            if f == 0:
                sim.spawn(recs)
                P.spawn(recs)
            sim.particle_step()
            P.step(w, DT, G, VS)
            _compare(sim, P, "floor frame %d" % f)
            overflowed += int(P.ctr[oracle.P_WRITES] > 500)
        assert overflowed > 0
        settled = (P.parts["value"] & N.P_SETTLED) != 0
        assert settled.sum() > 5000
    finally:
        gw.close()


def test_smallest_ring_wraps_in_one_spawn(small_scene):
    s = small_scene
    w = oracle.World.from_scene(s)
    gw, sim = _open(s, 10000)
    try:
        ring = 1 << 12
        sim.particles_init(ring)
        P = oracle.Particles(ring, spawn_budget=10000)
        recs = debris(s.dims, 9000, seed=3)
        for f in range(3):
            print("case: ring 4096 frame %d" % f, flush=True)
            sim.frame_begin()
            P.frame_begin()
            sim.spawn(recs[:7000])
            P.spawn(recs[:7000])
            sim.spawn(recs[7000:])  # crosses the budget: the tail is dropped
            P.spawn(recs[7000:])
            sim.particle_step()
            P.step(w, DT, G, VS)
            _compare(sim, P, "ring frame %d" % f)
    finally:
        gw.close()


def test_largest_ring_and_device_spawn(small_scene):
    import torch
    s = small_scene
    w = oracle.World.from_scene(s)
    gw, sim = _open(s)
    try:
        ring = 1 << 21
        sim.particles_init(ring)
        P = oracle.Particles(ring)
        recs = debris(s.dims, 60000, seed=11, speed=200.0)
        dev = torch.from_numpy(recs.view(np.uint8).copy()).cuda()
        for f in range(3):
            print("case: ring 2^21 frame %d" % f, flush=True)
            sim.frame_begin()
            P.frame_begin()
            sim.spawn_device(dev, len(recs))
            P.spawn(recs)
            sim.particle_step()
            P.step(w, DT, G, VS)
            _compare(sim, P, "2^21 frame %d" % f)
        assert int(P.ctr[oracle.P_ALIVE]) > 100000
    finally:
        gw.close()


def test_sim_step_stage_mask(small_scene):
    s = small_scene
    w = oracle.World.from_scene(s)
    gw, sim = _open(s)
    try:
        sim.particles_init(1 << 13)
        P = oracle.Particles(1 << 13)
        recs = debris(s.dims, 5000, seed=21)
        sim.spawn(recs)
        P.spawn(recs)
        for f in range(4):
            print("case: stage mask frame %d" % f, flush=True)
            sim.encode(N.STAGE_FRAME_BEGIN | N.STAGE_EDIT | N.STAGE_PARTICLES)
            P.frame_begin()
            P.step(w, DT, G, VS)
            _compare(sim, P, "encode frame %d" % f)
    finally:
        gw.close()


def test_particle_errors(small_scene):
    from pixels_amd import PxError
    gw, sim = _open(small_scene)
    try:
        with pytest.raises(PxError) as e:
            sim.particle_step()
        assert e.value.code == -4
        with pytest.raises(PxError):
            sim.encode(N.STAGE_PARTICLES)
        for bad in (3000, 1 << 11, 1 << 22):
            with pytest.raises(PxError) as e:
                sim.particles_init(bad)
            assert e.value.code == -1
        sim.particles_init(0)  # the reference default, 2^18
        assert len(sim.read_particles()) == 1 << 18
        with pytest.raises(PxError):
            gw.check(gw.lib.px_read_particles(gw.ctx, (1 << 18) - 1, 2, np.zeros(2, N.PARTICLE_DTYPE).ctypes.data))
        with pytest.raises(PxError):
            sim.encode(128)  # no such stage
    finally:
        gw.close()
