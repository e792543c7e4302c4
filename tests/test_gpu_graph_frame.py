"""A whole frame of px_sim_step stages (FRAME_BEGIN | SHOOT | EDIT | RESOLVE |
CHUNKS | PARTICLES | WRITES) recorded once with px_capture_sim_step and
replayed with px_graph_launch gives, frame after frame, the same world,
particle ring, chunk table and counters as the same stages called eagerly on
a second context: the frame runs with no host decision, allocation or wait.

The scene: a window with a block resting on it; the volley shatters the
window, the block becomes a falling chunk, lands and is written back, and
the debris deposits voxels, so every stage has work.  Between the replays the
test reads the device state back (eager copies and waits), and every replay
is compared.
"""
import numpy as np
import pytest

from pixels_amd import _native as N

pytestmark = pytest.mark.gpu

VS = np.float32(1.0 / 80.0)
DIMS = (128, 64, 96)
FRAME = (N.STAGE_FRAME_BEGIN | N.STAGE_SHOOT | N.STAGE_EDIT | N.STAGE_RESOLVE | N.STAGE_CHUNKS |
         N.STAGE_PARTICLES | N.STAGE_WRITES)


def _world():
    vox = np.zeros(DIMS[::-1], np.uint16)  # z, y, x
    vox[:, 0:4, :] = 1
    vox[20:70, 4:41, 60:62] = 9             # a window, top row y = 40
    vox[30:50, 41:49, 56:66] = 5            # a block resting on it
    vox[60:80, 4:12, 20:40] = 3             # something for the later shots to hit
    return vox


def _uniforms():
    u = np.zeros(N.PX_UNIFORM_FLOATS, np.float32)
    u[0:3] = DIMS
    u[3] = VS
    u[24] = np.float32(1.0 / 60.0)
    u[25] = np.float32(9.81) / VS
    u[28], u[29] = 2.0, 17.0
    u[40:44] = (6, 3.0, 5.0, 250.0)
    u[44:46] = (0.1, 5.0)
    return u


def _camera():
    cam = np.zeros(N.PX_CAMERA_FLOATS, np.float32)
    cam[32:35] = np.array((25.5, 30.5, 45.5), np.float32) * VS
    d = np.array((1.0, 0.02, -0.05))
    cam[36:39] = d / np.linalg.norm(d)
    return cam


def _open(scene):
    from pixels_amd import GpuWorld, Simulation
    lim = N.default_limits()
    gw = GpuWorld(scene, limits=lim)
    sim = Simulation(gw)
    gw.set_uniforms(_uniforms())
    sim.particles_init(1 << 14)
    sim.set_camera(_camera())
    return gw, sim


def _state(gw, sim):
    ctrl, args, _ = sim.read_ctrl()
    table = np.zeros(512 * 16, np.uint32)  # at most 511 chunk ids
    n = N.C.c_uint32(0)
    gw.check(gw.lib.px_read_chunks(gw.ctx, N.ptr(table), None, 0, N.C.byref(n)))
    return {"vox": sim.probe((0, 0, 0), DIMS), "ring": sim.read_particles(), "ctrl": ctrl, "args": args,
            "chunks": table, "flashes": sim.read_flashes()}


def test_frame_graph_replays_equal_eager_frames():
    from pixels_amd import Scene
    s = Scene.from_dense(_world())
    ga, sa = _open(s)
    gb, sb = _open(s)
    graph = None
    try:
        with pytest.raises(N.PxError) as e:  # no stage has run yet: nothing to record
            sb.capture(FRAME)
        assert e.value.code == -4  # PX_ERR_STATE
        # frame 0 eagerly on both (allocates every stage's buffers)
        sa.encode(FRAME)
        sb.encode(FRAME)
        graph = sb.capture(FRAME)
        falling = 0
        for f in range(1, 7):
            print("case: frame %d, eager vs graph replay" % f, flush=True)
            sa.encode(FRAME)
            graph.launch()
            a, b = _state(ga, sa), _state(gb, sb)
            for k in a:
                assert a[k].tobytes() == b[k].tobytes(), "frame %d: %s differs" % (f, k)
            falling = max(falling, int(a["ctrl"][N.CTRL["S_FALLING"]]) + int(a["ctrl"][N.CTRL["C_CHUNKS"]]))
        # every stage had work: the window went, a chunk fell, debris and writes
        assert not (a["vox"].reshape(_world().shape) == 9).any()
        assert falling > 0
        assert int(a["ctrl"][N.CTRL["C_SPAWNED"]]) > 0 or int(a["ctrl"][N.CTRL["C_PCURSOR"]]) > 0
    finally:
        if graph is not None:
            graph.close()
        ga.close()
        gb.close()
