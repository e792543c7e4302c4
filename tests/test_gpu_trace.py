"""GPU parity of the traversal kernels against the CPU oracle, bit-exact:
hit voxel key, face, t bits, step count and material for every ray.

Cases: the small seeded scene (512x256x512) with a 640x360 camera (primary,
the shadow-ray frame of shade(), shadow-kind batches), shoot rays
(traceSolid) and debris segments on the same scene, random small worlds
with glass, uniform bricks and rays from outside the volume, the empty
batch and decoded occupancy through probe.
"""
import numpy as np
import pytest

import oracle
from tests.conftest import rand_world

pytestmark = pytest.mark.gpu


def _assert_same(a, b, what):
    assert a.shape == b.shape
    bad = np.nonzero((a["key"] != b["key"]) | (a["face"] != b["face"]) | (a["steps"] != b["steps"]) |
                     (a["t"].view(np.uint32) != b["t"].view(np.uint32)) | (a["mat"] != b["mat"]) |
                     (a["value_hi"] != b["value_hi"]))[0]
    assert bad.size == 0, "%s: %d mismatches, first %s gpu=%s cpu=%s" % (what, bad.size, bad[:5], a[bad[:3]],
                                                                         b[bad[:3]])


@pytest.fixture(scope="module")
def small(small_scene):
    from pixels_amd import GpuWorld, Simulation
    s = small_scene
    print("case: small scene", s.dims, flush=True)
    gw = GpuWorld(s)
    sim = Simulation(gw)
    yield s, gw, sim, oracle.World.from_scene(s)
    gw.close()


def test_primary_rays_bit_exact(small):
    s, gw, sim, ow = small
    rays = s.camera_rays(640, 360, 90.0, seed=1)
    print("case: primary 640x360", flush=True)
    got = sim.trace(0, rays)
    ref = ow.trace(0, rays)
    _assert_same(got, ref, "primary")
    assert (ref["face"] != 0xFF).mean() > 0.3


def test_shadow_kind_bit_exact(small):
    s, gw, sim, ow = small
    rays = s.camera_rays(320, 180, 70.0, seed=5)
    print("case: shadow kind 320x180", flush=True)
    got = sim.trace(1, rays)
    ref = ow.trace(1, rays)
    _assert_same(got, ref, "shadow")


def test_frame_shadow_rays_bit_exact(small):
    import torch
    s, gw, sim, ow = small
    W, H = 320, 180
    rays = s.camera_rays(W, H, 90.0, seed=3)
    u, lights = s.frame_inputs(W, H, seed=3)
    gw.set_uniforms(u)
    gw.set_lights(lights)
    slots = sim.frame_slots()
    assert slots == 9
    print("case: frame 320x180 x 9 slots", flush=True)
    dr = torch.from_numpy(rays.view(np.uint8).copy()).cuda()
    dp = torch.zeros(W * H * 16, dtype=torch.uint8, device="cuda")
    ds = torch.zeros(W * H * slots * 16, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    nc = sim.trace_frame_device(dr, W * H, dp, ds, count=True)
    gw.sync()
    prim = dp.cpu().numpy().view(oracle.HIT_DTYPE)
    sh = ds.cpu().numpy().view(oracle.HIT_DTYPE)
    rp, rs, rnc = ow.frame(u, lights, rays)
    _assert_same(prim, rp, "frame primary")
    _assert_same(sh, rs, "frame shadow")
    assert nc == rnc
    assert nc > W * H  # several shadow rays per pixel


def test_shoot_and_segment_bit_exact(small):
    s, gw, sim, ow = small
    rng = np.random.default_rng(9)
    n = 20000
    rays = np.zeros(n, oracle.RAY_DTYPE)
    eye = np.array(s.camera(seed=9)[:3])
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["ro"] = eye
    rays["rd"] = d
    rays["max_t"] = 80.0 / (1 / 80.0)
    print("case: shoot 20000", flush=True)
    _assert_same(sim.trace(2, rays), ow.trace(2, rays), "shoot")
    segs = np.zeros(n, oracle.RAY_DTYPE)
    segs["ro"] = rng.uniform([0, 0, 0], s.dims, size=(n, 3))
    dd = rng.normal(size=(n, 3))
    dd /= np.linalg.norm(dd, axis=1, keepdims=True)
    segs["rd"] = dd * rng.uniform(1, 200, size=(n, 1))
    print("case: segment 20000", flush=True)
    _assert_same(sim.trace(3, segs), ow.trace(3, segs), "segment")
    mixed = np.concatenate([rays[:5000], segs[:5000]])
    mixed["flags"][:5000] = 2
    mixed["flags"][5000:] = 3
    print("case: mixed kinds", flush=True)
    _assert_same(sim.trace(4, mixed), ow.trace(-1, mixed), "mixed")


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_worlds_bit_exact(seed):
    from pixels_amd import GpuWorld, Scene, Simulation
    rng = np.random.default_rng(100 + seed)
    vox = rand_world(rng, dims=(96, 64, 80), fill=0.05, glass_ids=(9, 10, 11, 12), blobs=20)
    s = Scene.from_dense(vox)
    print("case: random world", seed, flush=True)
    gw = GpuWorld(s)
    sim = Simulation(gw)
    ow = oracle.World.from_scene(s)
    n = 30000
    rays = np.zeros(n, oracle.RAY_DTYPE)
    c = np.array(s.dims) / 2
    inside = rng.random(n) < 0.7
    ro = np.where(inside[:, None], rng.uniform([0, 0, 0], s.dims, size=(n, 3)),
                  c + rng.normal(size=(n, 3)) * 100)
    d = rng.normal(size=(n, 3))
    d[rng.random(n) < 0.05, 0] = 0.0
    d[rng.random(n) < 0.05, 1] = -0.0
    d[rng.random(n) < 0.02] = [0, 0, 1]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays["ro"] = ro
    rays["rd"] = d
    rays["max_t"] = np.where(rng.random(n) < 0.2, rng.uniform(0, 50, n), 1e9)
    for kind in (0, 1, 2):
        _assert_same(sim.trace(kind, rays), ow.trace(kind, rays), "random world kind %d" % kind)
    # occupancy through probe equals the dense world
    got = sim.probe((0, 0, 0), s.dims).reshape(vox.shape)
    assert np.array_equal(got, vox.astype(np.uint32))
    # a box partly outside the world reads air outside
    part = sim.probe((-4, -4, -4), (16, 16, 16)).reshape(16, 16, 16)
    assert np.all(part[:4] == 0) and np.array_equal(part[4:, 4:, 4:], vox[:12, :12, :12])
    gw.close()


def test_degenerate_directions_bit_exact():
    """Directions with NaN, zero and tiny components, origins on cell faces
    and inside glass, on a world whose last super cells are cut by its edge
    (136 x 40 x 72 voxels: 17 x 5 x 9 bricks).  A NaN component makes every
    comparison with its crossing time false and its normal zero; the
    transmittance of a camera ray through glass then reads a NaN cosine."""
    from pixels_amd import GpuWorld, Scene, Simulation
    rng = np.random.default_rng(77)
    dims = (136, 40, 72)
    vox = rand_world(rng, dims=dims, fill=0.06, glass_ids=(9, 10, 11, 12), blobs=16)
    s = Scene.from_dense(vox)
    print("case: degenerate directions", dims, flush=True)
    gw = GpuWorld(s)
    sim = Simulation(gw)
    ow = oracle.World.from_scene(s)
    n = 40000
    rays = np.zeros(n, oracle.RAY_DTYPE)
    d = np.asarray(dims, np.float32)
    ro = rng.uniform(-0.3, 1.3, (n, 3)).astype(np.float32) * d
    face = rng.random(n) < 0.3
    ro[face] = np.round(ro[face] * 2) / 2
    rd = rng.normal(size=(n, 3)).astype(np.float32)
    rd /= np.linalg.norm(rd, axis=1, keepdims=True)
    rd[rng.random((n, 3)) < 0.1] = 0.0
    rd[rng.random((n, 3)) < 0.05] = 3e-7
    rd[rng.random((n, 3)) < 0.08] = np.nan
    rd[(rd == 0).all(1)] = [0, 0, 1]
    rays["ro"] = ro
    rays["rd"] = rd
    rays["max_t"] = np.where(rng.random(n) < 0.3, rng.uniform(1, 300, n), 1e9)
    for kind in (0, 1):
        _assert_same(sim.trace(kind, rays), ow.trace(kind, rays), "degenerate kind %d" % kind)
    gw.close()


@pytest.mark.parametrize("n", [1, 63, 65, 257, 1000])
def test_ragged_batches_bit_exact(small, n):
    """Ray counts that leave partial wavefronts and workgroups: every kind of
    batch and the frame (whose shadow queue has one region per wavefront)."""
    import torch
    s, gw, sim, ow = small
    rays = s.camera_rays(200, 120, 80.0, seed=11)[:n]
    print("case: ragged n=%d" % n, flush=True)
    _assert_same(sim.trace(0, rays), ow.trace(0, rays), "primary n=%d" % n)
    _assert_same(sim.trace(1, rays), ow.trace(1, rays), "shadow n=%d" % n)
    u, lights = s.frame_inputs(200, 120, seed=11)
    gw.set_uniforms(u)
    gw.set_lights(lights)
    slots = sim.frame_slots()
    dr = torch.from_numpy(rays.view(np.uint8).copy()).cuda()
    dp = torch.zeros(n * 16, dtype=torch.uint8, device="cuda")
    ds = torch.zeros(n * slots * 16, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    nc = sim.trace_frame_device(dr, n, dp, ds, count=True)
    gw.sync()
    rp, rs, rnc = ow.frame(u, lights, rays)
    _assert_same(dp.cpu().numpy().view(oracle.HIT_DTYPE), rp, "frame primary n=%d" % n)
    _assert_same(ds.cpu().numpy().view(oracle.HIT_DTYPE), rs, "frame shadow n=%d" % n)
    assert nc == rnc


def test_empty_batch_and_bad_kind(small):
    from pixels_amd import PxError
    s, gw, sim, ow = small
    out = sim.trace(0, np.zeros(0, oracle.RAY_DTYPE))
    assert out.shape == (0,)
    with pytest.raises(PxError):
        gw.check(gw.lib.px_trace(gw.ctx, 9, None, 0, None))


def test_create_rejects_grid_beyond_the_uploaded_pool():
    """px_create checks every pool brick index of the grid, and every hole,
    against pool_used: a bad index would send the traversal's pool reads out
    of bounds."""
    import ctypes as C
    from pixels_amd import _native as N
    from pixels_amd.scene import Scene
    mats = np.ascontiguousarray(Scene.from_dense(np.zeros((8, 8, 8), np.uint16)).mats, np.float32)
    lib = N.load()
    pool = np.zeros(512, np.uint16)
    for grid, holes, ok in ((np.array([1, 0], np.uint32), None, True),
                            (np.array([2, 0], np.uint32), None, False),
                            (np.array([1, 0], np.uint32), np.array([1], np.uint32), False)):
        desc = N.PxWorldDesc()
        desc.dims[:] = (16, 8, 8)
        desc.grid = N.ptr(grid)
        desc.pool = N.ptr(pool, C.c_uint16)
        desc.pool_used = 1
        if holes is not None:
            desc.holes = N.ptr(holes)
            desc.n_holes = holes.size
        desc.mats = N.ptr(mats, C.c_float)
        ctx = C.c_void_p()
        rc = lib.px_create(C.byref(desc), C.byref(ctx))
        if ok:
            assert rc == 0
            lib.px_destroy(ctx)
        else:
            assert rc == -1, rc
