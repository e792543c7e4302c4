"""Pins the CPU oracle of the scene rasteriser (o_raster, apply() of
src/world/shapes.ts:334-414, primitive after primitive over the world):

* for a single primitive of every type, yawed or not, the voxels it sets are
  exactly those whose centre has SDF <= 0 (the brick and sub-block
  shortcuts are exact for these 1-Lipschitz SDFs), and the cell kinds follow
  compactCell;
* on the seeded scenes it equals the scene builder (pixels_amd/csrc/scene.cpp),
  an independent implementation that replays the primitives brick by brick.
"""
import numpy as np
import pytest

import oracle
from pixels_amd import _native as N
from tests.test_gpu_edit import kinds_of

DIMS = (48, 40, 56)


def _prim(t, p, mat=5, yaw=0.0, lo=None, hi=None):
    r = np.zeros(1, N.PRIM_DTYPE)
    r["type"] = t
    r["mat"] = mat
    r["p"][0, :len(p)] = p
    r["cy"], r["sy"] = (1.0, 0.0) if yaw == 0.0 else (np.cos(yaw), np.sin(yaw))
    r["lo"] = lo
    r["hi"] = hi
    return r


CASES = [
    _prim(N.PRIM_BOX, [20.3, 18.1, 25.7, 9.2, 5.5, 7.1], lo=[5, 5, 10], hi=[36, 32, 40]),
    _prim(N.PRIM_BOX, [24.0, 20.0, 28.0, 10.0, 6.0, 3.0], yaw=0.6, lo=[8, 13, 12], hi=[40, 27, 44]),
    _prim(N.PRIM_ROUNDBOX, [22.5, 19.5, 27.2, 11.0, 8.0, 9.0, 3.5], yaw=1.1, lo=[6, 10, 10], hi=[40, 30, 44]),
    _prim(N.PRIM_SPHERE, [23.4, 19.9, 28.1, 12.7], lo=[10, 7, 15], hi=[37, 33, 41]),
    _prim(N.PRIM_CYLINDER, [24.1, 20.2, 27.3, 6.5, 14.0, 1], yaw=0.3, lo=[17, 6, 20], hi=[31, 35, 35]),
    _prim(N.PRIM_CYLINDER, [24.1, 20.2, 27.3, 5.5, 16.0, 0], yaw=0.9, lo=[4, 14, 7], hi=[44, 26, 48]),
    _prim(N.PRIM_CAPSULE, [10.2, 8.7, 12.1, 38.4, 31.9, 44.6, 3.3], lo=[6, 5, 8], hi=[42, 36, 48]),
    _prim(N.PRIM_TORUS, [24.0, 20.0, 28.0, 12.0, 3.0, 2], yaw=0.4, lo=[8, 4, 12], hi=[40, 36, 44]),
    _prim(N.PRIM_AABB, [3.0, 0.0, 5.0, 44.0, 9.0, 50.0], lo=[3, 0, 5], hi=[44, 9, 50]),
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_single_primitive_is_exact(k):
    pr = CASES[k]
    vox, kind = oracle.raster(DIMS, pr)
    dx, dy, dz = DIMS
    ins = np.zeros((dz, dy, dx), bool)
    for z in range(dz):
        for y in range(dy):
            for x in range(dx):
                ins[z, y, x] = oracle.prim_sdf(pr, x + 0.5, y + 0.5, z + 0.5) <= 0.0
    assert ins.sum() > 50
    assert np.array_equal(vox != 0, ins)
    assert np.all(vox[ins] == 5)
    # kinds: empty / uniform / pool as compactCell leaves them
    b = vox.reshape(dz // 8, 8, dy // 8, 8, dx // 8, 8).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 512)
    exp = np.where(~(b != 0).any(1), 0, np.where((b == b[:, :1]).all(1), 1, 2))
    assert np.array_equal(kind, exp)


@pytest.mark.parametrize("kind_dims", [(0, (128, 64, 128)), (1, (128, 64, 96)), (2, (256, 96, 192))])
def test_equals_scene_builder(kind_dims):
    from pixels_amd.scene import Scene
    kind, dims = kind_dims
    s = Scene.build(kind, seed=5, dims=dims, n_prims=80)
    vox, kinds = oracle.raster(s.dims, s.prims())
    assert np.array_equal(vox, s.dense())
    assert np.array_equal(kinds, kinds_of(s.grid))
