"""GPU parity of the island pass (carve -> pane flood -> maint -> islands)
against the CPU oracle: decoded occupancy of every voxel, the canonical brick
encoding after maint (empty / uniform value / pool brick per cell), super-grid
counts, the free-list top, destroyed-voxel and bill counters, and the
(minimum voxel key, voxel count) of every unsupported component.

Cases: the seeded 512x256x512 scene with a 16-voxel structural floor and 64
impacts at primary-ray hit points (radii uniform 8-96), small worlds with
thin window panes (the flood), a pass with no impacts (labelling only), two
passes in a row, and an impact on uniform bricks beyond the pool capacity
(alloc failures are counted and leave those cells untouched)."""
import numpy as np
import pytest

import oracle
from tests.conftest import rand_world

pytestmark = pytest.mark.gpu


def kinds_of(grid):
    return np.where(grid == 0, 0, np.where(grid & 0x80000000, 1, 2)).astype(np.uint8)


def run_gpu(scene, impacts, pool_capacity=0):
    from pixels_amd import GpuWorld, Simulation
    gw = GpuWorld(scene, pool_capacity=pool_capacity)
    sim = Simulation(gw)
    if len(impacts):
        sim.push_impacts(impacts)
    sim.island_pass()
    keys, cnt, n = sim.island_labels()
    vox = sim.probe((0, 0, 0), scene.dims).reshape(scene.dims[2], scene.dims[1], scene.dims[0])
    grid, sgrid = gw.read_grid()
    ctrl, _, _ = sim.read_ctrl()
    return gw, sim, dict(keys=keys, cnt=cnt, n=n, vox=vox, grid=grid, sgrid=sgrid, ctrl=ctrl)


def check_against_oracle(scene, impacts, got, exact_free=True):
    from pixels_amd import CTRL
    dense = scene.dense()
    kind = kinds_of(scene.grid)
    ovox, okind, okeys, ocnt, on, stats = oracle.island_pass(scene.dims, dense, kind, scene.mats, impacts)
    assert np.array_equal(got["vox"], ovox.astype(np.uint32)), "occupancy differs"
    gk = kinds_of(got["grid"])
    assert np.array_equal(gk, okind), "brick encoding differs after maint"
    uni = got["grid"][gk == 1] & 0xFFFF
    # uniform values equal the decoded voxels
    bx, by, bz = scene.bricks
    g3 = got["grid"].reshape(bz, by, bx)
    zs, ys, xs = np.nonzero(g3 & 0x80000000)
    assert np.array_equal(uni, ovox[zs * 8, ys * 8, xs * 8].astype(np.uint32))
    # super grid counts follow the grid
    sx, sy, sz = scene.supers
    pad = np.zeros((sz * 4, sy * 4, sx * 4), np.uint32)
    pad[:bz, :by, :bx] = g3 != 0
    assert np.array_equal(pad.reshape(sz, 4, sy, 4, sx, 4).sum((1, 3, 5)).ravel(), got["sgrid"])
    ctrl = got["ctrl"]
    assert int(ctrl[CTRL["S_DESTROYED"]]) == int(stats[0])
    assert int(ctrl[CTRL["S_BILL"]]) == int(stats[1])
    assert got["n"] == on
    assert np.array_equal(got["keys"], okeys) and np.array_equal(got["cnt"], ocnt)
    return stats


@pytest.fixture(scope="module")
def floor_impacts(floor_scene):
    from pixels_amd import GpuWorld, Simulation
    from pixels_amd.world import impacts_from_hits
    s = floor_scene
    gw = GpuWorld(s)
    sim = Simulation(gw)
    rng = np.random.default_rng(8)
    rays = s.camera_rays(640, 360, 90.0, seed=3)
    pick = rays[rng.permutation(rays.shape[0])[:400]]
    hits = sim.trace(0, pick)
    gw.close()
    ok = hits["face"] != 0xFF
    pick, hits = pick[ok][:64], hits[ok][:64]
    r = rng.uniform(8, 96, pick.shape[0]).astype(np.float32)
    return impacts_from_hits(s.dims, pick, hits, r)


def test_island_pass_64_impacts_matches_oracle(floor_scene, floor_impacts):
    s = floor_scene
    assert floor_impacts.shape[0] == 64
    print("case: 64 impacts on", s.dims, flush=True)
    gw, sim, got = run_gpu(s, floor_impacts)
    stats = check_against_oracle(s, floor_impacts, got)
    assert stats[0] > 100000  # a lot was destroyed
    assert got["n"] > 0
    from pixels_amd import CTRL
    ncell_pool = int((kinds_of(got["grid"]) == 2).sum())
    assert gw.poolCapacity - int(got["ctrl"][CTRL["C_FREE_TOP"]]) == ncell_pool
    assert int(got["ctrl"][CTRL["S_ALLOC_FAIL"]]) == 0
    gw.close()


def test_island_pass_without_impacts(floor_scene):
    s = floor_scene
    print("case: labelling only", flush=True)
    gw, sim, got = run_gpu(s, np.zeros((0, 16), np.float32))
    okeys, ocnt, on = oracle.islands(s.dims, s.dense(), s.mats)
    assert got["n"] == on and np.array_equal(got["keys"], okeys) and np.array_equal(got["cnt"], ocnt)
    assert np.array_equal(got["grid"], s.grid)
    gw.close()


def _pane_world(seed):
    rng = np.random.default_rng(seed)
    dims = (96, 64, 80)
    vox = rand_world(rng, dims=dims, fill=0.01, glass_ids=(10, 11), blobs=6)
    vox[vox == 9] = 3
    vox[vox == 12] = 3
    # thin window panes (id 9) and glass table tops (id 12) in pool bricks
    for _ in range(6):
        z0, y0, x0 = rng.integers(0, [70, 50, 80])
        vox[z0:z0 + 10, y0:y0 + 14, x0:x0 + 2] = 9
        vox[z0 + 2:z0 + 8, y0:y0 + 1, x0 + 3:x0 + 15] = 12
    vox[:, 0:4, :] = 1  # structural floor
    return vox


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pane_flood_matches_oracle(seed):
    from pixels_amd import Scene
    vox = _pane_world(seed)
    s = Scene.from_dense(vox)
    rng = np.random.default_rng(50 + seed)
    panes = np.argwhere(vox == 9)
    n = 10
    pick = panes[rng.permutation(panes.shape[0])[:n]]
    imp = np.zeros((n, 16), np.float32)
    imp[:, 0] = pick[:, 2] + 0.5
    imp[:, 1] = pick[:, 1] + 0.5
    imp[:, 2] = pick[:, 0] + 0.5
    imp[:, 3] = rng.uniform(1.5, 6, n)
    imp[:, 4] = imp[:, 3]
    print("case: pane world", seed, flush=True)
    gw, sim, got = run_gpu(s, imp)
    stats = check_against_oracle(s, imp, got)
    assert stats[4] > 0  # the flood removed voxels beyond the carve
    gw.close()


def test_two_passes_in_a_row(floor_scene, floor_impacts):
    from pixels_amd import CTRL, GpuWorld, Simulation
    s = floor_scene
    gw = GpuWorld(s)
    sim = Simulation(gw)
    a, b = floor_impacts[:20], floor_impacts[20:40]
    print("case: two passes", flush=True)
    sim.push_impacts(a)
    sim.island_pass()
    sim.push_impacts(b)
    sim.island_pass()
    keys, cnt, n = sim.island_labels()
    vox = sim.probe((0, 0, 0), s.dims).reshape(s.dims[2], s.dims[1], s.dims[0])
    grid, _ = gw.read_grid()
    d0 = s.dense()
    k0 = kinds_of(s.grid)
    v1, k1, *_ = oracle.island_pass(s.dims, d0, k0, s.mats, a)
    v2, k2, okeys, ocnt, on, _ = oracle.island_pass(s.dims, v1, k1, s.mats, b)
    assert np.array_equal(vox, v2.astype(np.uint32))
    assert np.array_equal(kinds_of(grid), k2)
    assert n == on and np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt)
    ctrl, _, _ = sim.read_ctrl()
    assert int(ctrl[CTRL["S_SEARCHES"]]) == 2
    # traversal after the edits sees the edited world
    rays = s.camera_rays(160, 90, 90.0, seed=3)
    ow = oracle.World.from_scene(__import__("pixels_amd").Scene.from_dense(v2, s.mats))
    got = sim.trace(0, rays)
    ref = ow.trace(0, rays)
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
    gw.close()


def test_alloc_failure_is_counted():
    from pixels_amd import CTRL, Scene
    vox = np.zeros((64, 64, 64), np.uint16)
    vox[:, :32, :] = 3  # a slab of uniform bricks
    s = Scene.from_dense(vox)
    imp = np.zeros((1, 16), np.float32)
    imp[0, :3] = (32.0, 31.5, 32.0)
    imp[0, 3] = 20.0
    imp[0, 4] = 20.0
    print("case: alloc failure", flush=True)
    gw, sim, got = run_gpu(s, imp, pool_capacity=4)
    assert int(got["ctrl"][CTRL["S_ALLOC_FAIL"]]) > 0
    # only the 4 cells that got a pool brick were carved; the rest is intact
    removed = (vox != 0) & (got["vox"] == 0)
    cells = np.unique(np.stack(np.nonzero(removed), 1) // 8, axis=0)
    assert 0 < cells.shape[0] <= 4
    gw.close()


def test_sim_step_and_quiescence(floor_scene, floor_impacts):
    """px_sim_step(PX_STAGE_EDIT) runs the same pass as px_island_pass (and is
    a no-op without queued impacts); px_run_to_quiescence leaves no pending
    impacts; unknown stage bits are an error."""
    from pixels_amd import GpuWorld, PxError, Simulation
    s = floor_scene
    gw = GpuWorld(s)
    sim = Simulation(gw)
    try:
        print("case: sim_step", flush=True)
        sim.encode()  # nothing queued: no-op
        gw.sync()
        grid0, _ = gw.read_grid()
        assert np.array_equal(grid0, s.grid)
        sim.push_impacts(floor_impacts[:32])
        sim.encode()
        keys, cnt, n = sim.island_labels()
        vox = sim.probe((0, 0, 0), s.dims).reshape(s.dims[2], s.dims[1], s.dims[0])
        grid, sgrid = gw.read_grid()
        ctrl, _, _ = sim.read_ctrl()
        check_against_oracle(s, floor_impacts[:32], dict(keys=keys, cnt=cnt, n=n, vox=vox, grid=grid, sgrid=sgrid,
                                                          ctrl=ctrl))
        sim.run_to_quiescence()  # nothing pending any more
        grid2, _ = gw.read_grid()
        assert np.array_equal(grid2, grid)
        with pytest.raises(PxError):
            sim.encode(1 << 10)  # not a stage bit
    finally:
        gw.close()


def test_walk_bound_refuses_labels(floor_scene, floor_impacts):
    """A union-find walk that passes its step bound stops (every kernel still
    exits) and the labels are refused; the bound is forced tiny from the
    host, so the well-formed walks of the 64-impact pass exceed it."""
    from pixels_amd import GpuWorld, PxError, Simulation, default_limits
    lim = default_limits()
    lim.uf_walk_cap = 1
    gw = GpuWorld(floor_scene, limits=lim)
    try:
        sim = Simulation(gw)
        sim.push_impacts(floor_impacts)
        print("case: walk bound 1", flush=True)
        with pytest.raises(PxError):
            sim.run_to_quiescence()
        assert "union-find" in gw.lib.px_last_error(gw.ctx).decode()
        with pytest.raises(PxError):
            sim.island_labels()
    finally:
        gw.close()
    # the default bound: the same pass is accepted
    gw = GpuWorld(floor_scene)
    try:
        sim = Simulation(gw)
        sim.push_impacts(floor_impacts)
        sim.run_to_quiescence()
        keys, cnt, n = sim.island_labels()
        assert n > 0
    finally:
        gw.close()
