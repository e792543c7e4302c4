import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def has_gpu() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def small_scene():
    from pixels_amd.scene import SMALL, Scene
    return Scene.build(SMALL, seed=1)


@pytest.fixture(scope="session")
def floor_scene():
    from pixels_amd.scene import SMALL_FLOOR, Scene
    return Scene.build(SMALL_FLOOR, seed=3)


def rand_world(rng: np.random.Generator, dims=(64, 32, 48), fill=0.08, glass_ids=(9, 10), blobs=12,
               uniform_blocks=4):
    """Small random world: scattered voxels, a few blobs and some full bricks."""
    dx, dy, dz = dims
    vox = np.zeros((dz, dy, dx), np.uint16)
    m = rng.random((dz, dy, dx)) < fill
    vals = rng.integers(1, 17, size=(dz, dy, dx)).astype(np.uint16)
    vox[m] = vals[m]
    for _ in range(blobs):
        c = rng.integers(0, [dz, dy, dx])
        r = rng.integers(2, 7)
        z0, y0, x0 = np.maximum(c - r, 0)
        z1, y1, x1 = c + r
        vox[z0:z1, y0:y1, x0:x1] = rng.choice(list(glass_ids) + [2, 3, 5])
    for _ in range(uniform_blocks):
        bz, by, bx = rng.integers(0, [dz // 8, dy // 8, dx // 8])
        vox[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8] = rng.choice([2, 3, 9])
    return vox
