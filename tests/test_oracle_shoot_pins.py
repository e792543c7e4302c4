"""Independent pins of the oracle's shoot stage (oracle/px_oracle.c o_shoot,
restating sim.wgsl:371-403 with pushImpact :203-211, takeVoxel :105-124,
shardEffects :261-271, dropRubbleAbove :274-284 and pushFlood :250-258):

* a pellet fired along +x at the top row of a 2-voxel-thick pane hits the
  pane's first voxel with the -x face; the impact centre is the voxel centre
  moved min(0.3 r, 2) into the wall; the record holds r, fr, impulse, the
  normal and the direction; the flash sits 4 voxels out along the normal,
  in metres;
* pellets traced on the same world hit the same voxel: four pellets queue four
  impacts, only the first takes the voxel, so one flood entry; the rubble
  column above it tumbles (one DEPOSIT record per rubble voxel, bottom to
  top, each 0.5 voxel above the last) before the shard record;
* the flood entry empties the whole pane in the following edit pass even when
  the impact carves nothing (a tiny radius), and the cells the takes touched
  are maintained;
* spread pellets: unit directions within the cone the spread allows, and each
  impact's voxel is the one o_trace_solid finds on the brick map along the
  impact's direction (pins the dense-world tracing of o_shoot);
* two planted mistakes (the centre moved out of the wall, the shot's flood
  entries ignored) each fail one of these checks.
"""
import numpy as np
import pytest

import oracle

VS = np.float32(1.0 / 80.0)
PANE, RUBBLE = 9, 2 | 0x80


@pytest.fixture(scope="module")
def mats():
    from pixels_amd.scene import Scene
    return Scene.from_dense(np.zeros((8, 8, 8), np.uint16)).mats


def _kinds(vox):
    dz, dy, dx = vox.shape
    b = vox.reshape(dz // 8, 8, dy // 8, 8, dx // 8, 8).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 512)
    k = np.where((b != 0).any(1), 2, 0).astype(np.uint8)
    uni = (b == b[:, :1]).all(1) & (b[:, 0] != 0)
    k[uni] = 1
    return k


def _world():
    vox = np.zeros((64, 64, 64), np.uint16)  # z, y, x
    vox[:, 0:4, :] = 1                        # structural floor
    vox[10:51, 4:41, 30:32] = PANE            # 2 voxels thick, top row y = 40
    vox[30, 41:46, 30] = RUBBLE               # 5 rubble voxels on the struck column
    vox[20:26, 4:20, 8:14] = 5                # an ordinary block
    return vox


def _uniforms(pellets, r, fr, impulse, spread=0.0, seed=0.0, frame=7.0, time_x=3.25):
    u = np.zeros(52, np.float32)
    u[0:3] = (64, 64, 64)
    u[3] = VS
    u[28], u[29] = time_x, frame
    u[40:44] = (pellets, r, fr, impulse)
    u[44:46] = (spread, seed)
    return u


def _camera(eye_vox, look):
    cam = np.zeros(64, np.float32)
    cam[32:35] = np.asarray(eye_vox, np.float32) * VS
    d = np.asarray(look, np.float64)
    cam[36:39] = d / np.linalg.norm(d)
    return cam


def _shoot(mats, vox, u, cam, lib=None):
    return oracle.shoot((64, 64, 64), vox, _kinds(vox), mats, u, cam, lib=lib)


def test_straight_shot_closed_forms(mats):
    vox = _world()
    r, fr, imp = 6.0, 4.0, 250.0
    s = _shoot(mats, vox, _uniforms(4, r, fr, imp), _camera((10.5, 40.5, 30.5), (1, 0, 0)))
    assert s.n_impacts == 4 and s.stats[2] == 4
    off = min(0.3 * r, 2.0)
    for k in range(4):  # every pellet hit voxel (30, 40, 30) through its -x face
        im = s.impacts[k]
        assert np.allclose(im[0:3], (30.5 + off, 40.5, 30.5), atol=1e-5)
        assert tuple(im[3:6]) == (np.float32(r), np.float32(fr), np.float32(imp))
        assert tuple(im[6:9]) == (-1.0, 0.0, 0.0)
        assert np.allclose(im[9:12], (1, 0, 0), atol=1e-7)
    # one take: the struck voxel and the rubble column above it
    assert s.n_flood == 1 and tuple(s.flood[0]) == (30 + 64 * (40 + 64 * 30), PANE)
    assert s.vox[30, 40, 30] == 0 and (s.vox[30, 41:46, 30] == 0).all()
    assert (s.vox[30, 39, 30] == PANE) and (s.vox[30, 40, 31] == PANE)
    assert s.stats[0] == 1  # destroyed: the pane voxel (rubble drops are not billed)
    sp = s.spawns
    assert s.spawn_calls in (5, 6)
    assert np.allclose(sp["pos"][:5], [(30.5, 41.5 + k, 30.5) for k in range(5)])
    assert ((sp["value"][:5] & 0xFFFF) == RUBBLE).all() and ((sp["value"][:5] & 0x20000) != 0).all()
    if s.spawn_calls == 6:
        assert tuple(sp["pos"][5]) == (30.5, 40.5, 30.5) and (sp["value"][5] & 0xFF) == PANE
    # flash of pellet 0, record 0, metres
    assert np.allclose(s.flashes[0, 0:3], np.array((30.5 + off - 4.0, 40.5, 30.5)) * VS, atol=1e-6)
    assert s.flashes[0, 3] == np.float32(r) * VS and s.flashes[0, 4] == np.float32(3.25)
    assert s.flash_ctr == 1 and not s.flashes[1:].any()
    # the takes dirtied the struck cell and the cell of the rubble above it
    cells = set(np.nonzero(s.dirty)[0].tolist())
    assert cells == {3 + 8 * (5 + 8 * 3)}


def test_shot_floods_the_whole_pane(mats):
    vox = _world()
    u = _uniforms(1, 0.1, 0.1, 100.0)  # an impact that carves nothing
    s = _shoot(mats, vox, u, _camera((10.5, 40.5, 30.5), (1, 0, 0)))
    e = oracle.edit_pass((64, 64, 64), s.vox, s.kind, mats, s.impacts[:s.n_impacts], vs=VS, frame=7.0,
                         flood=s.flood, dirty=s.dirty)
    assert not (e.vox == PANE).any()
    assert int(e.stats[4]) == int((vox == PANE).sum()) - 1  # flooded = the pane minus the struck voxel
    assert (e.vox[20:26, 4:20, 8:14] == 5).all() and (e.vox[:, 0:4, :] == 1).all()
    # without the flood entry the pane stays
    e2 = oracle.edit_pass((64, 64, 64), s.vox, s.kind, mats, s.impacts[:s.n_impacts], vs=VS, frame=7.0)
    assert int((e2.vox == PANE).sum()) == int((vox == PANE).sum()) - 1


def test_spread_pellets_trace_the_brick_map(mats):
    from pixels_amd.scene import Scene
    vox = _world()
    u = _uniforms(16, 3.0, 2.0, 100.0, spread=0.3, seed=12345.0)
    cam = _camera((12.5, 30.5, 30.5), (1, 0.1, 0.05))
    s = _shoot(mats, vox, u, cam)
    assert s.n_impacts >= 10
    ow = oracle.World.from_scene(Scene.from_dense(vox))
    look = cam[36:39].astype(np.float64)
    for k in range(min(s.n_impacts, 64)):
        im = s.impacts[k]
        d = im[9:12].astype(np.float64)
        assert abs(np.linalg.norm(d) - 1.0) < 1e-6
        # |dir - look| <= |(rnd3 - 0.5) * 2 * spread| / |look + that| <= 0.3 * sqrt(3) / (1 - 0.3 * sqrt(3))
        assert np.linalg.norm(d - look) < 0.3 * np.sqrt(3) / (1 - 0.3 * np.sqrt(3)) + 1e-6
        ray = np.zeros(1, oracle.RAY_DTYPE)
        ray["ro"] = cam[32:35] / VS
        ray["rd"] = im[9:12]
        ray["max_t"] = np.float32(80.0) / VS
        h = ow.trace(2, ray)[0]
        v = (im[0:3] + im[6:9] * np.float32(min(0.3 * 3.0, 2.0)) - 0.5).round().astype(np.int64)
        assert h["key"] == v[0] + 64 * (v[1] + 64 * v[2])
    dirs = s.impacts[:s.n_impacts, 9:12]
    assert len({tuple(d) for d in dirs}) > 8  # the pellets spread


def test_no_pellets_no_effects(mats):
    vox = _world()
    s = _shoot(mats, vox, _uniforms(0, 6.0, 4.0, 100.0), _camera((10.5, 40.5, 30.5), (1, 0, 0)))
    assert s.n_impacts == 0 and s.n_flood == 0 and s.spawn_calls == 0 and s.flash_ctr == 0
    assert np.array_equal(s.vox, vox)


@pytest.mark.parametrize("plant", ["O_PLANT_SHOOT_CENTRE_SIGN", "O_PLANT_SHOT_FLOOD_SKIPPED"])
def test_planted_mistakes_fail_a_pin(mats, plant, tmp_path):
    lib = oracle.load_variant(str(tmp_path / ("lib_%s.so" % plant)), [plant])
    vox = _world()
    if plant == "O_PLANT_SHOOT_CENTRE_SIGN":
        r = 6.0
        s = _shoot(mats, vox, _uniforms(1, r, 4.0, 100.0), _camera((10.5, 40.5, 30.5), (1, 0, 0)), lib=lib)
        assert not np.allclose(s.impacts[0, 0:3], (30.5 + min(0.3 * r, 2.0), 40.5, 30.5), atol=1e-5)
    else:
        s = _shoot(mats, vox, _uniforms(1, 0.1, 0.1, 100.0), _camera((10.5, 40.5, 30.5), (1, 0, 0)), lib=lib)
        e = oracle.edit_pass((64, 64, 64), s.vox, s.kind, mats, s.impacts[:s.n_impacts], vs=VS, frame=7.0,
                             flood=s.flood, dirty=s.dirty, lib=lib)
        assert (e.vox == PANE).any()
