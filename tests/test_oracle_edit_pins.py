"""Independent pins of the oracle's edit pass (oracle/px_oracle.c o_edit_pass,
restating sim.wgsl:229-248, 274-284, 414-618 and the island search :620-882).

* Impact jobs + alloc: on worlds of materials with softness <= 1 (whose
  removal radius max(r, fr) * 1.2 the job box R covers), the carved set equals
  removes() evaluated by brute force over every voxel of the world, uniform
  and pool bricks alike.
* Pane flood: the removed pane set equals the union of the
  scipy.ndimage.label components (6-connected, one pane id at a time, pool
  voxels) that hold a carved voxel, minus the carved voxels.
* Three planted mistakes (a job box without the x1.2, reach() without the
  fragile radius, a flood that crosses pane ids) are built into copies of the
  oracle, and each fails one of the pins above.
* Effects: closed-form rubble stacks (dropRubbleAbove), the debris fractions
  keep * E.w and 0.6 * E.w, velocity bounds and means, lives, and the record
  order (carved voxels, then flooded ones, each in cell / local-index order).
* Seeded search: the seeded labels equal the scipy components of the final
  world that are floating and hold a seed; a floating piece no damage touches
  stays out; the node budget keeps the longest label-order prefix within it.
"""
import numpy as np
import pytest
from scipy import ndimage

import oracle

SOFT_LE_1 = [2, 3, 4, 5, 8, 10, 11, 13, 14, 15, 16]  # softness <= 1, not panes, not structural


@pytest.fixture(scope="module")
def mats():
    from pixels_amd.scene import Scene
    return Scene.from_dense(np.zeros((8, 8, 8), np.uint16)).mats


def kinds(vox, all_pool=False):
    dz, dy, dx = vox.shape
    b = vox.reshape(dz // 8, 8, dy // 8, 8, dx // 8, 8).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 512)
    k = np.where((b != 0).any(1), 2, 0).astype(np.uint8)
    if not all_pool:
        uni = (b == b[:, :1]).all(1) & (b[:, 0] != 0)
        k[uni] = 1
    return k


def impact(c, r, fr, impulse=0.0, n=(0, 0, 0), d=(0, 0, 0)):
    im = np.zeros(16, np.float32)
    im[0:3], im[3], im[4], im[5] = c, r, fr, impulse
    im[6:9], im[9:12] = n, d
    return im


def brute_removed(mats, vox, impacts):
    """removes() of every impact over every non-air voxel (numpy distance
    prefilter at 1.68 max(r, fr) + 1, then the oracle's per-voxel predicate)."""
    out = np.zeros(vox.shape, bool)
    zz, yy, xx = np.nonzero(vox)
    for im in impacts:
        lim = 1.68 * max(im[3], im[4]) + 2.0
        near = ((xx + 0.5 - im[0]) ** 2 + (yy + 0.5 - im[1]) ** 2 + (zz + 0.5 - im[2]) ** 2) < lim * lim
        for z, y, x in zip(zz[near], yy[near], xx[near]):
            if not out[z, y, x] and oracle.removes(im, mats, int(x), int(y), int(z), int(vox[z, y, x])):
                out[z, y, x] = True
    return out


def job_world(seed):
    """Boxes of soft <= 1 materials, several fully solid bricks (stored
    uniform), fragile uniform bricks near the impacts, no panes."""
    rng = np.random.default_rng(seed)
    dims = (96, 64, 80)
    vox = np.zeros(dims[::-1], np.uint16)
    for _ in range(30):
        c = rng.integers(0, [80, 64, 96])
        h = rng.integers(3, 14, 3)
        lo, hi = np.maximum(c - h, 0), c + h
        vox[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = rng.choice(SOFT_LE_1)
    for _ in range(12):  # brick-aligned blocks: uniform bricks, fragile ones among them
        b = rng.integers(0, [10, 8, 12]) * 8
        vox[b[0]:b[0] + 16, b[1]:b[1] + 8, b[2]:b[2] + 8] = rng.choice([10, 11, 14, 16, 3, 5])
    imps = []
    for _ in range(14):
        c = rng.uniform([4, 4, 4], [92, 60, 76]).astype(np.float32)
        r = np.float32(rng.uniform(2.0, 9.0))
        fr = np.float32(r * rng.uniform(1.0, 2.5))
        imps.append(impact(c, r, fr))
    return vox, np.array(imps)


def edit(mats, vox, kind, imps, lib=None, **kw):
    return oracle.edit_pass(vox.shape[::-1], vox, kind, mats, imps, lib=lib, **kw)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_carve_set_equals_brute_force_removes(mats, seed):
    vox, imps = job_world(seed)
    kind = kinds(vox)
    assert (kind == 1).sum() > 20
    exp = brute_removed(mats, vox, imps)
    r = edit(mats, vox, kind, imps)
    got = (vox != 0) & (r.vox == 0)
    assert exp.sum() > 500
    assert np.array_equal(got, exp), "carved %d, brute force %d" % (got.sum(), exp.sum())
    assert int(r.stats[0]) == int(exp.sum())
    assert int(r.stats[4]) == 0  # no panes: no flood


def pane_world(seed):
    rng = np.random.default_rng(seed)
    dims = (96, 64, 80)
    vox = np.zeros(dims[::-1], np.uint16)
    vox[:, 0:4, :] = 1
    for _ in range(10):  # windows (id 9) with glass shelves (id 12) touching them
        z0, y0, x0 = rng.integers(2, [64, 44, 76])
        vox[z0:z0 + 14, y0:y0 + 16, x0:x0 + 2] = 9
        vox[z0 + 3:z0 + 10, y0 + 5, x0 + 2:x0 + 16] = 12
        vox[z0 + 3:z0 + 10, y0 + 9, x0 - 6:x0] = 12
    for _ in range(8):
        c = rng.integers(0, [80, 64, 96])
        vox[c[0]:c[0] + 6, c[1]:c[1] + 6, c[2]:c[2] + 6] = rng.choice([2, 3, 5])
    # impacts at pane voxels
    zz, yy, xx = np.nonzero((vox == 9) | (vox == 12))
    pick = rng.choice(zz.size, 6, replace=False)
    imps = np.array([impact(np.array([xx[i] + 0.5, yy[i] + 0.5, zz[i] + 0.5], np.float32), np.float32(2.5),
                            np.float32(3.0)) for i in pick])
    return vox, imps


def expected_flood(mats, vox, carved):
    """Same-id 6-connected components of pane voxels (scipy), pool voxels
    everywhere, that hold a carved voxel; minus the carved voxels."""
    flood = np.zeros(vox.shape, bool)
    st = ndimage.generate_binary_structure(3, 1)
    for pid in (9, 12):
        lab, n = ndimage.label(vox == pid, structure=st)
        hit = np.unique(lab[carved & (vox == pid)])
        hit = hit[hit != 0]
        flood |= np.isin(lab, hit)
    return flood & ~carved


@pytest.mark.parametrize("seed", [4, 5])
def test_flood_equals_hit_components(mats, seed):
    vox, imps = pane_world(seed)
    kind = kinds(vox, all_pool=True)
    carved = brute_removed(mats, vox, imps)
    flood = expected_flood(mats, vox, carved)
    r = edit(mats, vox, kind, imps)
    exp_vox = np.where(carved | flood, 0, vox)
    assert flood.sum() > 200
    assert np.array_equal(r.vox, exp_vox)
    assert int(r.stats[4]) == int(flood.sum())
    # both pane ids were struck, and some panes stay whole
    assert ((vox == 9) & flood).any() and ((vox == 12) & flood).any()
    assert ((vox == 9) & (r.vox == 9)).any() or ((vox == 12) & (r.vox == 12)).any()


@pytest.mark.parametrize("plant", ["O_PLANT_AABB_NO_1_2", "O_PLANT_REACH_NO_FR", "O_PLANT_FLOOD_ANY_PANE"])
def test_planted_mistakes_fail_a_pin(mats, plant, tmp_path):
    lib = oracle.load_variant(str(tmp_path / ("lib_%s.so" % plant)), [plant])
    failed = False
    for seed in (1, 2, 3):
        vox, imps = job_world(seed)
        exp = brute_removed(mats, vox, imps)
        r = edit(mats, vox, kinds(vox), imps, lib=lib)
        failed |= not np.array_equal((vox != 0) & (r.vox == 0), exp)
    for seed in (4, 5):
        vox, imps = pane_world(seed)
        carved = brute_removed(mats, vox, imps)
        flood = expected_flood(mats, vox, carved)
        r = edit(mats, vox, kinds(vox, all_pool=True), imps, lib=lib)
        failed |= not np.array_equal(r.vox, np.where(carved | flood, 0, vox))
    assert failed, "%s passes every pin" % plant


# ---- effects -------------------------------------------------------------

VS = np.float32(1.0 / 80.0)


def test_rubble_stacks(mats):
    """dropRubbleAbove: rubble columns on carved voxels tumble as DEPOSIT
    debris bottom to top, at most 12, only rubble of pool bricks; a non-rubble
    voxel ends a column; rubble on an uncarved voxel stays."""
    dims = (32, 48, 32)
    vox = np.zeros(dims[::-1], np.uint16)
    vox[:, 0:4, :] = 1
    vox[14:18, 4:8, 14:18] = 5           # a small block (softness 1), carved away
    vox[15, 8:13, 15] = 3 | 0x80         # 5 rubble voxels on it (softness 0.4: out of the carve)
    vox[16, 8:28, 16] = 8 | 0x80         # 20 rubble voxels: 12 drop
    vox[14, 8:10, 14] = 3 | 0x80         # 2 rubble, then stone, then rubble
    vox[14, 10, 14] = 2
    vox[14, 11:13, 14] = 4 | 0x80
    vox[2, 4:9, 2] = 3 | 0x80            # rubble on the floor, far away: stays
    kind = kinds(vox, all_pool=True)
    imps = np.array([impact((16.0, 5.0, 16.0), 5.0, 5.0, impulse=0.0)])
    r = edit(mats, vox, kind, imps, vs=VS, frame=3.0)
    carved = (vox != 0) & (r.vox == 0) & ((vox & 0x80) == 0)
    assert carved[14:18, 4:8, 14:18].all()
    assert np.array_equal(r.vox[15, 8:13, 15], [0] * 5)
    assert np.array_equal(r.vox[16, 8:20, 16], [0] * 12) and np.array_equal(r.vox[16, 20:28, 16], [8 | 0x80] * 8)
    assert np.array_equal(r.vox[14, 8:13, 14], [0, 0, 2, 4 | 0x80, 4 | 0x80])
    assert np.array_equal(r.vox[2, 4:9, 2], [3 | 0x80] * 5)
    dep = r.spawns[(r.spawns["value"] & oracle.PARTICLE_DEPOSIT) != 0]
    rub = dep[(dep["value"] & 0x80) != 0]
    assert len(rub) == 5 + 12 + 2
    # each column bottom to top, right after ... its carved voxel; zero vertical speed, |vx|, |vz| <= 0.2 / vs
    for x, z, n in ((15, 15, 5), (16, 16, 12), (14, 14, 2)):
        col = rub[(rub["pos"][:, 0] == x + 0.5) & (rub["pos"][:, 2] == z + 0.5)]
        assert len(col) == n
        assert np.array_equal(col["pos"][:, 1], 8.5 + np.arange(n))
    assert np.all(rub["vel"][:, 1] == 0.0)
    assert np.all(np.abs(rub["vel"][:, [0, 2]]) <= np.float32(0.2) / VS)
    assert np.all((rub["life"] >= 6.0) & (rub["life"] <= 8.0))


def test_debris_fractions_velocities_and_order(mats):
    """Carve debris: every carved voxel of a debris-1.0 material spawns once
    when keep = 1; the fraction is keep * E.w otherwise; zero impulse gives
    zero velocity; the mean radial speed is 0.7 * E[s] = 0.525 impulse and the
    mean vertical jitter 0.15 impulse.  Shard debris: a 0.6 * E.w fraction,
    speeds within (0.8, 0.4, 0.8) / vs.  Records come carved voxels first,
    then flooded ones, each in (cell, local index) order."""
    dims = (128, 96, 128)
    vox = np.zeros(dims[::-1], np.uint16)
    vox[:, 0:4, :] = 1
    vox[10:118, 4:90, 10:118] = 14                   # debris 1.0
    kind = kinds(vox, all_pool=True)
    # r = 6: keep = min(1, 30000 / (4.19 r^3)) = 1
    im1 = impact((40.0, 40.0, 40.0), 6.0, 6.0, impulse=0.0)
    r = edit(mats, vox, kind, np.array([im1]), vs=VS, frame=2.0)
    carved = int(((vox != 0) & (r.vox == 0)).sum())
    assert len(r.spawns) == r.spawn_calls == carved
    assert np.all(r.spawns["vel"] == 0.0)
    assert np.all((r.spawns["life"] >= 5.0) & (r.spawns["life"] <= 9.0))
    assert np.all((r.spawns["value"] & 0xff) == 14)
    # r = 40: keep = 30000 / (4.19 * 64000) = 0.1119
    I = np.float32(300.0)
    im2 = impact((64.0, 48.0, 64.0), 40.0, 40.0, impulse=I)
    r = edit(mats, vox, kind, np.array([im2]), vs=VS, frame=2.0)
    carved = int(((vox != 0) & (r.vox == 0)).sum())
    keep = 30000.0 / (4.19 * 40.0 ** 3)
    n = len(r.spawns)
    sd = np.sqrt(carved * keep * (1 - keep))
    assert abs(n - carved * keep) < 5 * sd, (n, carved * keep)
    d = r.spawns["pos"] - np.array([64.0, 48.0, 64.0], np.float32)
    dirn = d / np.linalg.norm(d, axis=1, keepdims=True)
    radial = (r.spawns["vel"] * dirn).sum(1)
    assert abs(radial.mean() / I - 0.525) < 0.03 + 0.15 * abs(dirn[:, 1].mean())
    jit_y = r.spawns["vel"][:, 1] - 0.7 * dirn[:, 1] * (0.3 + 0.9 * 0.5) * I
    assert abs(jit_y.mean() / I - 0.15) < 0.05
    assert np.all(np.linalg.norm(r.spawns["vel"], axis=1) <= (0.7 * 1.2 + 0.5) * I)
    # order: (cell, local index) increasing
    p = r.spawns["pos"].astype(np.int64)
    bx, by = dims[0] // 8, dims[1] // 8
    cell = p[:, 0] // 8 + bx * (p[:, 1] // 8 + by * (p[:, 2] // 8))
    li = (p[:, 0] & 7) | ((p[:, 1] & 7) << 3) | ((p[:, 2] & 7) << 6)
    assert np.all(np.diff(cell * 512 + li) > 0)


def test_shard_debris_and_group_order(mats):
    vox, imps = pane_world(4)
    kind = kinds(vox, all_pool=True)
    carved = brute_removed(mats, vox, imps)
    flood = expected_flood(mats, vox, carved)
    r = edit(mats, vox, kind, imps, vs=VS, frame=7.0)
    p = r.spawns["pos"].astype(np.int64)
    in_carved = carved[p[:, 2], p[:, 1], p[:, 0]]
    in_flood = flood[p[:, 2], p[:, 1], p[:, 0]]
    assert np.all(in_carved | in_flood)
    k = np.nonzero(in_flood)[0]
    assert k.size > 0 and np.all(in_carved[:k[0]]) and np.all(in_flood[k[0]:])
    sh = r.spawns[k]
    # ids 9 (debris 0.5) and 12 (debris 0.8): fractions 0.3 and 0.48 of the flooded voxels
    for pid, frac in ((9, 0.6 * 0.5), (12, 0.6 * 0.8)):
        nf = int((flood & (vox == pid)).sum())
        ns = int(((sh["value"] & 0xff) == pid).sum())
        assert abs(ns - nf * frac) < 5 * np.sqrt(nf * frac * (1 - frac)) + 1, (pid, ns, nf * frac)
    v = sh["vel"] * VS
    assert np.all(np.abs(v[:, [0, 2]]) <= 0.8 + 1e-6) and np.all((v[:, 1] >= 0) & (v[:, 1] <= 0.4 + 1e-6))
    assert np.all((sh["life"] >= 4.0) & (sh["life"] <= 8.0))


# ---- seeded search ---------------------------------------------------------

def node_voxels(mats, vox):
    m = mats.reshape(256, 20)
    return (vox != 0) & ((vox & 0x80) == 0) & ~(m[vox & 0xff, 18] > 0.5)


def expected_seeded(mats, vox0, vox1):
    """Floating components (scipy, 6-connected node voxels of the final
    world; anchored by y == 0 or a structural neighbour) that hold a seed:
    a final node voxel next to a removed voxel, not of the removed voxel's id
    when that one is a pane."""
    m = mats.reshape(256, 20)
    st = ndimage.generate_binary_structure(3, 1)
    node = node_voxels(mats, vox1)
    lab, n = ndimage.label(node, structure=st)
    removed = (vox0 != 0) & (vox1 == 0)
    seed = np.zeros(vox0.shape, bool)
    pane0 = m[vox0 & 0xff, 17] > 1.5
    for ax in range(3):
        for sgn in (1, -1):
            src = np.roll(removed, sgn, axis=ax)
            srcv = np.roll(vox0, sgn, axis=ax)
            srcp = np.roll(pane0, sgn, axis=ax)
            edge = [slice(None)] * 3
            edge[ax] = 0 if sgn == 1 else -1
            src[tuple(edge)] = False
            seed |= src & node & ~(srcp & ((srcv & 0xff) == (vox1 & 0xff)))
    struct = (vox1 != 0) & ((vox1 & 0x80) == 0) & (m[vox1 & 0xff, 18] > 0.5)
    anch = np.zeros(n + 1, bool)
    anch[np.unique(lab[:, 0, :])] = True
    touch = ndimage.binary_dilation(struct, structure=st) & node
    anch[np.unique(lab[touch])] = True
    seeded = np.zeros(n + 1, bool)
    seeded[np.unique(lab[seed])] = True
    keys, counts = [], []
    flat = lab.ravel()
    first = np.full(n + 1, np.iinfo(np.int64).max, np.int64)
    idx = np.nonzero(flat)[0]
    np.minimum.at(first, flat[idx], idx)
    cnt = np.bincount(flat, minlength=n + 1)
    for c in range(1, n + 1):
        if seeded[c] and not anch[c]:
            keys.append(first[c])
            counts.append(cnt[c])
    o = np.argsort(keys)
    return np.array(keys, np.int64)[o], np.array(counts, np.int64)[o]


def seeded_world():
    dims = (96, 64, 96)
    vox = np.zeros(dims[::-1], np.uint16)
    vox[:, 0:4, :] = 1                       # structural floor
    vox[40:44, 4:40, 40:44] = 2              # a pillar on the floor ...
    vox[30:54, 40:44, 30:54] = 3             # ... carrying a table top
    vox[60:70, 30:40, 60:70] = 5             # a floating block no impact touches
    vox[10:14, 4:30, 10:14] = 8              # a second pillar ...
    vox[8:16, 30:34, 8:16] = 13              # ... with its own top
    vox[70:80, 20:22, 10:30] = 9             # a pane on a post
    vox[74:76, 4:20, 10:12] = 4
    return vox


def test_seeded_equals_scipy_seeded_components(mats):
    vox = seeded_world()
    kind = kinds(vox, all_pool=True)
    # cut both pillars through and the pane's post
    imps = np.array([impact((42.0, 20.0, 42.0), 6.0, 6.0), impact((12.0, 15.0, 12.0), 8.0, 8.0),
                     impact((11.0, 12.0, 75.0), 6.0, 6.0)])
    r_all = edit(mats, vox, kind, imps, mode=0)
    r_seed = edit(mats, vox, kind, imps, mode=1)
    assert np.array_equal(r_all.vox, r_seed.vox)
    ek, ec = expected_seeded(mats, vox, r_seed.vox)
    assert np.array_equal(r_seed.keys, ek) and np.array_equal(r_seed.counts, ec)
    # the untouched floating block: labelled by the global form only
    block_key = 60 + 96 * (30 + 64 * 60)
    assert block_key in set(r_all.keys.tolist()) and block_key not in set(r_seed.keys.tolist())
    assert len(r_seed.keys) >= 2  # the table top and the second top fell
    assert int(r_seed.stats[5]) > 0 and int(r_seed.stats[6]) == 0


def test_node_budget_keeps_label_order_prefix(mats):
    vox = seeded_world()
    kind = kinds(vox, all_pool=True)
    imps = np.array([impact((42.0, 20.0, 42.0), 6.0, 6.0), impact((12.0, 15.0, 12.0), 8.0, 8.0),
                     impact((11.0, 12.0, 75.0), 6.0, 6.0)])
    full = edit(mats, vox, kind, imps, mode=1)
    cs = np.cumsum(full.counts)
    for budget in (0, int(cs[0]), int(cs[0]) + 1, int(cs[-1]) - 1, int(cs[-1])):
        r = edit(mats, vox, kind, imps, mode=1, node_budget=budget)
        k = int(np.searchsorted(cs, budget, side="right"))
        assert np.array_equal(r.keys, full.keys[:k]) and np.array_equal(r.counts, full.counts[:k])
        assert int(r.stats[6]) == len(full.keys) - k


def test_landing_seeds_join_the_search(mats):
    """A pending seed (pushSeed under a landed chunk) makes its component
    searched even without impacts."""
    vox = seeded_world()
    kind = kinds(vox, all_pool=True)
    key = 62 + 96 * (35 + 64 * 62)  # inside the floating block
    r0 = edit(mats, vox, kind, np.zeros((0, 16), np.float32), mode=1)
    r1 = edit(mats, vox, kind, np.zeros((0, 16), np.float32), mode=1, seed_keys=np.array([key], np.uint32))
    assert r0.n == 0 and r1.n == 1 and int(r1.keys[0]) == 60 + 96 * (30 + 64 * 60) and int(r1.counts[0]) == 1000
