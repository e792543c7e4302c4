"""Pins the CPU oracle of the voxel-write stage (o_apply_writes in
oracle/px_oracle.c: writeRequest / allocBricks / writeApply / maint,
sim.wgsl:94-146, 459-483, 1178-1203, 1219-1274) with hand-computed answers:
a write into an empty cell gets a brick, a taken voxel stacks the write one
or two voxels higher and a fourth write is dropped, uniform bricks and the
top of the world count as taken / no brick, the stack stops at a cell
without a brick, the first max_alloc distinct cells in list order are the
ones served, and a brick filled with one value collapses to uniform."""
import numpy as np

import oracle

DIMS = (16, 16, 16)


def _key(x, y, z, dims=DIMS):
    return x + dims[0] * (y + dims[1] * z)


def _world():
    vox = np.zeros((DIMS[2], DIMS[1], DIMS[0]), np.uint16)
    kind = np.zeros(8, np.uint8)  # 2x2x2 cells, x fastest
    return vox, kind


def _cell(x, y, z):
    return (x >> 3) + 2 * ((y >> 3) + 2 * (z >> 3))


def test_write_into_empty_cell_allocates():
    vox, kind = _world()
    v1, k1, st = oracle.apply_writes(DIMS, vox, kind, [[_key(3, 2, 1), 0x1285]])
    assert v1[1, 2, 3] == 0x1285 and np.count_nonzero(v1) == 1
    assert k1[_cell(3, 2, 1)] == 2 and np.count_nonzero(k1) == 1
    assert list(st) == [1, 1, 0, 1]


def test_stacking_and_drop():
    vox, kind = _world()
    w = [[_key(5, 3, 5), 10 + i] for i in range(4)]
    v1, k1, st = oracle.apply_writes(DIMS, vox, kind, w)
    assert v1[5, 3, 5] == 10 and v1[5, 4, 5] == 11 and v1[5, 5, 5] == 12
    assert np.count_nonzero(v1) == 3  # the fourth write found p, p+y, p+2y taken
    assert st[0] == 3 and st[1] == 1


def test_uniform_cell_is_taken_and_world_top_is_no_brick():
    vox, kind = _world()
    vox[0:8, 0:8, 0:8] = 7
    kind[0] = 1
    # lands in the uniform brick: p, p+y taken (uniform), p+2y = y 8 is in an empty cell -> stop
    v1, k1, st = oracle.apply_writes(DIMS, vox, kind, [[_key(2, 6, 2), 99]])
    assert np.array_equal(v1, vox) and np.array_equal(k1, kind) and st[0] == 0
    # the cell is not requested: only a write's first target asks for a brick
    assert st[1] == 0
    # top of the world: y = 15 taken, y = 16 is outside -> dropped
    vox2, kind2 = _world()
    vox2[3, 15, 3] = 5
    kind2[_cell(3, 15, 3)] = 2
    v2, k2, st2 = oracle.apply_writes(DIMS, vox2, kind2, [[_key(3, 15, 3), 6]])
    assert st2[0] == 0 and v2[3, 15, 3] == 5


def test_stack_crosses_into_allocated_cell():
    vox, kind = _world()
    vox[2, 7, 2] = 4
    kind[_cell(2, 7, 2)] = 2
    # first write targets (2,7,2): taken -> (2,8,2) in the upper cell, which has no brick
    # unless another write's first target is there; the second write requests it first
    w = [[_key(2, 7, 2), 20], [_key(6, 9, 2), 21]]
    v1, k1, st = oracle.apply_writes(DIMS, vox, kind, w)
    assert v1[2, 8, 2] == 20 and v1[2, 9, 6] == 21
    assert st[1] == 1 and st[0] == 2


def test_first_max_alloc_cells_in_list_order():
    vox, kind = _world()
    w = [[_key(1, 1, 1), 1], [_key(9, 1, 1), 2], [_key(2, 2, 2), 3], [_key(1, 9, 9), 4]]
    v1, k1, st = oracle.apply_writes(DIMS, vox, kind, w, max_alloc=1)
    assert v1[1, 1, 1] == 1 and v1[2, 2, 2] == 3
    assert v1[1, 1, 9] == 0 and v1[9, 9, 1] == 0
    assert st[1] == 1 and st[2] == 2
    v2, k2, st2 = oracle.apply_writes(DIMS, vox, kind, w, free_top=2)
    assert st2[1] == 2 and st2[2] == 1 and v2[9, 9, 1] == 0 and v2[1, 1, 9] == 2


def test_full_brick_collapses_to_uniform():
    vox, kind = _world()
    w = [[_key(8 + (l & 7), 8 + ((l >> 3) & 7), (l >> 6)), 0x33] for l in range(512)]
    v1, k1, st = oracle.apply_writes(DIMS, vox, kind, w)
    assert k1[_cell(8, 8, 0)] == 1 and np.all(v1[0:8, 8:16, 8:16] == 0x33)
    assert st[0] == 512
