"""Scene builder: the reference encoding (VoxelWorld.ts:4-13) is produced
canonically and the super-grid counts follow GpuWorld.ts:90-100."""
import numpy as np

from pixels_amd.scene import LARGE, SMALL, Scene


def _check_canonical(s):
    g = s.grid
    pool_cells = np.nonzero((g != 0) & ((g & 0x80000000) == 0))[0]
    idx = (g[pool_cells] - 1).astype(np.int64)
    assert np.array_equal(np.sort(idx), np.arange(s.pool_used))
    bricks = s.pool.reshape(-1, 512)
    # every pool brick holds at least two distinct values (compactCell)
    assert np.all(bricks.min(1) != bricks.max(1))
    # uniform entries carry a non-zero value
    uni = g[(g & 0x80000000) != 0] & 0xFFFF
    assert np.all(uni != 0)
    # super-grid counts
    bx, by, bz = s.bricks
    occ = (g != 0).reshape(bz, by, bx)
    sx, sy, sz = s.supers
    pad = np.zeros((sz * 4, sy * 4, sx * 4), np.uint32)
    pad[:bz, :by, :bx] = occ
    cnt = pad.reshape(sz, 4, sy, 4, sx, 4).sum((1, 3, 5)).ravel()
    assert np.array_equal(cnt, s.super_grid)


def test_dense_round_trip():
    rng = np.random.default_rng(3)
    vox = np.zeros((24, 16, 32), np.uint16)
    vox[rng.random(vox.shape) < 0.1] = 5
    vox[0:8, 0:8, 8:16] = 7
    s = Scene.from_dense(vox)
    assert np.array_equal(s.dense(), vox)
    _check_canonical(s)


def test_small_scene_properties():
    s = Scene.build(SMALL, seed=1)
    assert s.dims == (512, 256, 512)
    assert s.n_prims == 200
    _check_canonical(s)
    d = s.dense()
    assert int((d != 0).sum()) == s.solid
    ids = np.unique(d[d != 0] & 0xFF)
    assert ids.min() >= 1 and ids.max() <= 16
    # deterministic from the seed
    s2 = Scene.build(SMALL, seed=1)
    assert np.array_equal(s.grid, s2.grid) and np.array_equal(s.pool, s2.pool)


def test_large_scene_properties():
    s = Scene.build(LARGE, seed=2)
    assert s.dims == (1280, 352, 960)
    assert s.n_prims == 2002
    _check_canonical(s)
    # structural slabs: the bottom and top 16 voxel rows are solid material 1
    bx, by, bz = s.bricks
    g = s.grid.reshape(bz, by, bx)
    assert np.all(g[:, 0:2, :] == (0x80000000 | 1))
    assert np.all(g[:, by - 2:, :] == (0x80000000 | 1))
    assert 0.1 < s.fill < 0.5
    assert 0.03 < s.glass / s.solid < 0.2


def test_camera_rays_are_unit_and_tiled():
    s = Scene.build(SMALL, seed=1)
    r = s.camera_rays(64, 32, 90.0, seed=2, tile=8)
    n = np.linalg.norm(r["rd"].astype(np.float64), axis=1)
    assert np.allclose(n, 1.0, atol=1e-6)
    assert np.all(r["max_t"] == np.float32(1e9))
    # all rays share the eye, which sits in air
    assert np.all(r["ro"] == r["ro"][0])
    e = np.floor(r["ro"][0]).astype(int)
    assert s.dense()[e[2], e[1], e[0]] == 0
