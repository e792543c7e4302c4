"""GPU scene rasteriser (px_create_from_prims, apply() of
src/world/shapes.ts:334-414) against the oracle's restatement (o_raster) and
the host scene builder: the grid (pool bricks numbered in cell order), every
pool brick, the super-grid counts, and the decoded voxels, bit for bit.
Cases: one primitive of every type (yawed, thin, tiny), the seeded 512x256x512
scenes with and without the floor slab, the 1280x352x960 scene of 2002
primitives, and an empty list."""
import numpy as np
import pytest

import oracle
from pixels_amd import _native as N
from tests.test_gpu_edit import kinds_of
from tests.test_raster_oracle import CASES, DIMS

pytestmark = pytest.mark.gpu


def _gpu(prims, dims, mats):
    from pixels_amd import GpuWorld
    return GpuWorld(None, prims=prims, dims=dims, mats=mats)


def _pool(gw, used):
    return gw.read_pool(0, used).view(np.uint16).reshape(-1) if used else np.zeros(0, np.uint16)


def test_single_primitives_match_oracle(small_scene):
    from pixels_amd import Simulation
    for k, pr in enumerate(CASES):
        print("case: primitive %d type %d" % (k, int(pr["type"][0])), flush=True)
        gw = _gpu(pr, DIMS, small_scene.mats)
        try:
            vox, kind = oracle.raster(DIMS, pr)
            got = Simulation(gw).probe((0, 0, 0), DIMS).reshape(DIMS[2], DIMS[1], DIMS[0])
            assert np.array_equal(got, vox.astype(np.uint32)), k
            grid, _ = gw.read_grid()
            assert np.array_equal(kinds_of(grid), kind), k
        finally:
            gw.close()


@pytest.mark.parametrize("which", ["small", "floor", "large"])
def test_seeded_scenes_equal_host_builder(which, small_scene, floor_scene):
    import time
    from pixels_amd.scene import LARGE, Scene
    s = {"small": small_scene, "floor": floor_scene}.get(which) or Scene.build(LARGE, seed=2)
    prims = s.prims()
    print("case: scene %s %s, %d primitives" % (which, s.dims, len(prims)), flush=True)
    t0 = time.time()
    gw = _gpu(prims, s.dims, s.mats)
    build_s = time.time() - t0
    try:
        grid, sup = gw.read_grid()
        assert np.array_equal(grid, s.grid)
        assert np.array_equal(sup, s.super_grid)
        assert np.array_equal(_pool(gw, s.pool_used), s.pool)
        assert gw.poolCapacity == max(s.pool_used + 98304, 2 * s.pool_used)
        print("GPU build of %s: %.3f s wall (with context creation)" % (which, build_s), flush=True)
    finally:
        gw.close()


def test_oracle_on_a_mid_scene():
    from pixels_amd import Simulation
    from pixels_amd.scene import LARGE, Scene
    s = Scene.build(LARGE, seed=7, dims=(256, 96, 192), n_prims=150)
    gw = _gpu(s.prims(), s.dims, s.mats)
    try:
        vox, kind = oracle.raster(s.dims, s.prims())
        got = Simulation(gw).probe((0, 0, 0), s.dims).reshape(s.dims[2], s.dims[1], s.dims[0])
        assert np.array_equal(got, vox.astype(np.uint32))
        grid, _ = gw.read_grid()
        assert np.array_equal(kinds_of(grid), kind)
    finally:
        gw.close()


def test_empty_list(small_scene):
    gw = _gpu(np.zeros(0, N.PRIM_DTYPE), (64, 32, 64), small_scene.mats)
    try:
        grid, sup = gw.read_grid()
        assert not grid.any() and not sup.any()
    finally:
        gw.close()
