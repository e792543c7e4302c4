"""Pins the oracle's shading (oracle/px_shade.c: fsMain's colour, raymarch.wgsl:
167-450, materialColor render_common.wgsl:49-140) where the mathematics fixes
the answer:

* the elementary functions (sin, cos, exp, exp2, log2, pow, atan2) are
  correctly rounded against numpy's float64 results (exp and pow above the
  2^-126 flush), on dense sweeps, special values included;
* hsv() against colorsys, vnoise at lattice points against hashf;
* the density texture against a numpy count of opaque voxels per brick, and
  the trilinear clamp-to-edge sampler (coarse AO) against
  scipy.ndimage.map_coordinates (order 1, nearest);
* the sky at the zenith (exactly the zenith colour) and in the sun direction
  (+ sunColor * 40.35);
* debug views 1, 3, 4 and the miss colour against the hits of march();
* an emissive voxel is albedo * emission whatever the lights;
* a pixel in the sun's shadow equals the same pixel with the sun off, bit for
  bit; a light beyond its radius changes nothing;
* fine AO beside a wall: the closed form 0.35 + 0.65 (1 - 2/3 fu);
* a glass pane seen head-on scales what lies behind by its transmittance
  (1 - F)(0.4 + 0.6 tint) g and adds the Fresnel glow.
"""
import colorsys
import ctypes as C
import math

import numpy as np
import pytest

import oracle

VS = 1.0 / 80.0


# ---- the gallery world (shared with the GPU parity tests) -----------------
def palette():
    """Material table: the built-in entries plus ids 20..33 with patterns
    0..13, 40 emissive, 41 mirror, 42 glossy reflective, 43 / 44 glass."""
    from pixels_amd.scene import Scene
    m = Scene.from_dense(np.zeros((8, 8, 8), np.uint16)).mats.copy().reshape(256, 20)
    scales = [0.35, 0.05, 0.05, 0.3, 0.05, 0.05, 0.4, 0.02, 0.02, 0.01, 0.05, 0.08, 1.0, 0.05]
    for k in range(14):
        e = m[20 + k]
        e[:] = 0.0
        e[0:3] = [0.62, 0.41, 0.3]
        e[4:7] = [0.2, 0.52, 0.71]
        e[7] = k
        e[8:12] = [scales[k], 0.06, 0.5, 40.0]
        e[16] = 1.0
    e = m[40]
    e[:] = 0.0
    e[0:4] = [1.0, 0.85, 0.6, 6.0]
    e[8:12] = [0.1, 0.0, 0.0, 16.0]
    e = m[41]
    e[:] = 0.0
    e[0:3] = [0.08, 0.09, 0.1]
    e[8:12] = [0.1, 0.0, 1.0, 300.0]
    e[12] = 0.85
    e = m[42]
    e[:] = 0.0
    e[0:3] = [0.78, 0.8, 0.82]
    e[4:8] = [0.55, 0.58, 0.62, 10.0]
    e[8:12] = [0.05, 0.03, 1.0, 120.0]
    e[12] = 0.45
    for k, (tint, g) in enumerate([((0.78, 0.87, 0.91), 0.82), ((0.18, 0.48, 0.23), 0.7)]):
        e = m[43 + k]
        e[:] = 0.0
        e[0:3] = tint
        e[8:12] = [0.1, 0.08, 1.0, 200.0]
        e[13] = g
    return m.reshape(-1)


def gallery():
    """96x48x96: floor, one block per pattern on a circle, shade-byte voxels,
    an emissive block, a mirror wall, a glossy block, two glass panes and an
    overhang over the camera (open sky above)."""
    from pixels_amd.scene import Scene
    dims = (96, 48, 96)
    vox = np.zeros((dims[2], dims[1], dims[0]), np.uint16)
    vox[:, :4, :] = 1
    for k in range(14):
        a = 2.0 * math.pi * k / 14.0
        cx, cz = int(48 + 34 * math.cos(a)), int(48 + 34 * math.sin(a))
        vox[cz - 4:cz + 4, 4:12 + k, cx - 4:cx + 4] = 20 + k
    vox[38:42, 4:20, 24:28] = 5 | (37 << 8)
    vox[38:42, 20:24, 24:28] = 20 | (200 << 8)
    vox[60:64, 4:24, 66:70] = 40
    vox[22:32, 4:30, 78] = 41
    vox[64:70, 4:14, 36:42] = 42
    vox[54:70, 4:20, 56] = 43
    vox[34, 4:20, 40:56] = 44
    vox[42:54, 26:28, 42:54] = 3
    return Scene.from_dense(vox, palette())


def panorama(w=192, h=96, cam=(48.3, 14.6, 47.7)):
    """Equirectangular rays from cam: every direction of the sphere."""
    j, i = np.mgrid[0:h, 0:w]
    az = 2.0 * np.pi * (i + 0.5) / w
    el = np.pi * (j + 0.5) / h - np.pi / 2.0
    d = np.stack([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)], -1).reshape(-1, 3)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(d.shape[0], oracle.RAY_DTYPE)
    rays["ro"] = np.asarray(cam, np.float32)
    rays["rd"] = d.astype(np.float32)
    rays["max_t"] = np.float32(1e9)
    return rays


def uniforms(dims, night=0.3, reflections=1.0, lamp_shadows=1.0, debug=0.0, sun=(0.4, 0.7, 0.3), sun_w=1.0,
             ao=0.6, time=12.5):
    u = np.zeros(52, np.float32)
    u[0:3] = dims
    u[3] = VS
    s = np.asarray(sun, np.float64)
    u[4:7] = s / np.linalg.norm(s)
    u[7] = sun_w
    u[8:11] = [1.0, 0.92, 0.85]
    u[12:16] = [0.3, 0.31, 0.36, ao]
    u[20:24] = [reflections, lamp_shadows, 1.2, debug]
    u[24:28] = [1.0 / 60.0, 9.81 / VS, 262144.0, night]
    u[28] = time
    return u


def lights(entries=None):
    """[(pos voxels, radius voxels, colour, shadow flag)] -> the light buffer."""
    entries = entries if entries is not None else [
        ((30.0, 20.0, 30.0), 70.0, (1.0, 0.8, 0.6), 1.0),
        ((70.0, 10.0, 62.0), 50.0, (0.4, 0.6, 1.0), 1.0),
        ((50.0, 36.0, 80.0), 90.0, (1.0, 0.3, 0.5), 0.0),
        ((20.0, 8.0, 70.0), 40.0, (0.9, 0.9, 0.7), 1.0),
    ]
    out = np.zeros((1 + 2 * len(entries)) * 4, np.float32)
    out[0] = len(entries)
    for i, (p, r, c, sh) in enumerate(entries):
        out[(1 + 2 * i) * 4:(1 + 2 * i) * 4 + 4] = [p[0] * VS, p[1] * VS, p[2] * VS, r * VS]
        out[(2 + 2 * i) * 4:(2 + 2 * i) * 4 + 4] = [c[0], c[1], c[2], sh]
    return out


def f32bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


# ---- elementary functions -------------------------------------------------
def _ulps(got, exact):
    got = np.float64(np.float32(got))
    sp = np.float64(np.spacing(np.float32(abs(exact))))
    return abs(got - exact) / sp


FUNCS = {
    "sin": (np.sin, lambda r: np.concatenate([r.uniform(-7, 7, 4000), r.uniform(-3000, 3000, 4000)])),
    "cos": (np.cos, lambda r: np.concatenate([r.uniform(-7, 7, 4000), r.uniform(-3000, 3000, 4000)])),
    "exp2": (np.exp2, lambda r: r.uniform(-125, 127, 8000)),
    "exp": (np.exp, lambda r: r.uniform(-87, 88, 8000)),
    "log2": (np.log2, lambda r: np.exp(r.uniform(-85, 85, 8000))),
}


@pytest.mark.parametrize("name", sorted(FUNCS))
def test_function_correctly_rounded(name):
    ref, gen = FUNCS[name]
    xs = gen(np.random.default_rng(7)).astype(np.float32)
    worst = 0.0
    for x in xs:
        e = ref(np.float64(x))
        if name == "log2" and abs(e) < 1e-3:
            continue
        worst = max(worst, _ulps(oracle.fmath(name, x), e))
    assert worst <= 0.5 + 1e-6, worst


def test_pow_and_atan2_correctly_rounded():
    rng = np.random.default_rng(8)
    worst = 0.0
    for b, e in zip(rng.uniform(0, 1, 6000).astype(np.float32), rng.uniform(0.5, 1500, 6000).astype(np.float32)):
        ex = np.float64(b) ** np.float64(e)
        if ex > 2.0 ** -126:
            worst = max(worst, _ulps(oracle.fmath("pow", b, e), ex))
    assert worst <= 0.5 + 1e-6, worst
    worst = 0.0
    for y, x in zip(rng.normal(size=6000).astype(np.float32), rng.normal(size=6000).astype(np.float32)):
        worst = max(worst, _ulps(oracle.fmath("atan2", y, x), np.arctan2(np.float64(y), np.float64(x))))
    assert worst <= 0.5 + 1e-6, worst


def test_function_special_values():
    f = oracle.fmath
    assert f("pow", 0.0, 2.0) == 0.0 and f("pow", 0.0, 0.0) == 1.0 and f("pow", 1.0, 1500.0) == 1.0
    assert f("pow", 0.25, 0.5) == 0.5 and f("exp2", 10.0) == 1024.0 and f("log2", 0.125) == -3.0
    assert f("exp", 0.0) == 1.0 and f("exp", -200.0) == 0.0
    assert f("sin", 0.0) == 0.0 and f("cos", 0.0) == 1.0
    assert math.isnan(f("sin", float("inf"))) and math.isnan(f("pow", -1.0, 2.0))
    assert f("atan2", 0.0, -1.0) == np.float32(np.pi) and f("atan2", -0.0, -1.0) == -np.float32(np.pi)
    assert f("atan2", 1.0, 0.0) == np.float32(np.pi / 2) and f("atan2", 0.0, 1.0) == 0.0


# ---- noise, hsv, density, sampler -----------------------------------------
def test_hsv_matches_colorsys():
    """Screen pattern (11) colour = hsv(h, 0.85, v): checked through the
    oracle's materialColor on a material without noise."""
    m = np.zeros(256 * 20, np.float32)
    m[5 * 20 + 7] = 11.0          # pattern 11
    m[5 * 20 + 8] = 1.0           # scale 1 m
    u = uniforms((64, 64, 64), time=0.0)
    rng = np.random.default_rng(3)
    for _ in range(40):
        vp = rng.integers(0, 4000, 3).astype(np.int32)
        c = oracle.material_color(m, u, 5, vp)
        wp = (vp.astype(np.float32) + np.float32(0.5)) * np.float32(VS)
        row, col = np.floor(wp[1]), np.floor(wp[0] + wp[2])
        # hash2f(row + spin, col) with spin = floor(col * 1.7) at time 0
        sym = _hash2f(np.float32(row + np.floor(np.float32(col) * np.float32(1.7))), np.float32(col))
        h = (sym * np.float32(3.7)) % 1.0
        ref = np.array(colorsys.hsv_to_rgb(float(h), 0.85, 0.6 + 0.4 * float(sym)))
        if (wp[1] % 1.0) < 0.08:
            ref = ref * 0.15
        assert np.allclose(c, ref, atol=2e-6), (vp, c, ref)


def _hashu(x):
    x = int(x) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7feb352d) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846ca68b) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def _hash2f(x, y):
    a = (int(np.floor(x)) + 100000) & 0xFFFFFFFF
    b = (int(np.floor(y)) + 7777) & 0xFFFFFFFF
    return np.float32(_hashu(((a * 0x27d4eb2d) & 0xFFFFFFFF) ^ _hashu(b)) & 0xFFFF) / np.float32(65535.0)


def test_noise_pattern_at_lattice_points_is_the_hash():
    """Pattern 13 (leaf) = mix(A, B, vnoise(wp / s)); at a lattice point the
    noise is hashf of the point (u = 0)."""
    m = np.zeros(256 * 20, np.float32)
    m[7 * 20 + 0:7 * 20 + 3] = [0.0, 0.0, 0.0]
    m[7 * 20 + 4:7 * 20 + 7] = [1.0, 1.0, 1.0]
    m[7 * 20 + 7] = 13.0
    m[7 * 20 + 8] = 1.0 / 128.0   # with 1/64 m voxels, wp / s = 2 vp + 1 exactly
    u = uniforms((64, 64, 64))
    u[3] = 1.0 / 64.0
    for vp in [(0, 0, 0), (3, 9, 17), (100, 7, 2), (-5, 4, 8)]:
        c = oracle.material_color(m, u, 7, np.array(vp, np.int32))
        p = [2 * v + 1 for v in vp]
        h = np.float32(oracle.load().o_hash3i(*p) & 0xFFFFFF) / np.float32(16777215.0)
        assert c[0] == h and c[1] == h and c[2] == h, (vp, c, h)


def test_density_counts_opaque_voxels():
    s = gallery()
    ow = oracle.World.from_scene(s)
    d = ow.density()
    vox = s.dense()
    dz, dy, dx = vox.shape
    glass = s.mats.reshape(256, 20)[:, 13] > 0
    op = (vox != 0) & ~glass[vox & 0xFF]
    cnt = op.reshape(dz // 8, 8, dy // 8, 8, dx // 8, 8).sum(axis=(1, 3, 5))
    exp = np.floor(cnt * 255.0 / 512.0 + 0.5).astype(np.uint8)
    assert np.array_equal(d, exp.reshape(-1))
    assert d.min() == 0 and d.max() > 128 and len(np.unique(d)) > 10


def test_coarse_ao_sampler_is_trilinear_clamp_to_edge():
    """Debug view 6 (coarse AO x fine AO) on floor cells with no opaque voxel
    around them (fine AO exactly 1) and AO strength 1: the view is
    clamp(1 - sum w_k sample(P + N d_k), 0.12, 1) with a trilinear,
    edge-clamped sample of the density texture (scipy map_coordinates)."""
    from scipy.ndimage import map_coordinates
    s = gallery()
    ow = oracle.World.from_scene(s)
    d = ow.density().astype(np.float64) / 255.0
    bd = [v // 8 for v in s.dims]
    tex = d.reshape(bd[2], bd[1], bd[0])
    vox = s.dense()
    glass = s.mats.reshape(256, 20)[:, 13] > 0
    opaque = (vox != 0) & ~glass[vox & 0xFF]
    u = uniforms(s.dims, debug=6.0, ao=1.0)
    rays = panorama(128, 64)
    img = ow.render(u, lights([]), rays, threads=4)
    checked = 0
    for k in range(rays.shape[0]):
        hit, _, _, _, _ = ow.march_ex(u, rays["ro"][k], rays["rd"][k])
        if hit["face"] != 2:      # top faces (N = +y)
            continue
        x, y, z = hit["key"] % 96, (hit["key"] // 96) % 48, hit["key"] // (96 * 48)
        if not (1 <= x < 95 and 1 <= z < 95) or opaque[z - 1:z + 2, y + 1, x - 1:x + 2].any():
            continue
        P = rays["ro"][k].astype(np.float64) + rays["rd"][k].astype(np.float64) * float(hit["t"])
        occ = 0.0
        for dist, wt in ((10.0, 0.55), (22.0, 0.35), (44.0, 0.3)):
            q = (P + np.array([0.0, dist, 0.0])) / np.asarray(s.dims, np.float64) * np.asarray(bd, np.float64) - 0.5
            occ += wt * map_coordinates(tex, [[q[2]], [q[1]], [q[0]]], order=1, mode="nearest")[0]
        coarse = min(max(1.0 - occ, 0.12), 1.0)
        assert abs(float(img[k, 0]) - coarse) < 2e-5, (k, img[k, 0], coarse)
        checked += 1
    assert checked > 100


# ---- sky -------------------------------------------------------------------
def test_sky_zenith_and_sun():
    u = uniforms((64, 64, 64), night=0.0, sun=(0.2, -1.0, 0.1))
    c = oracle.sky_color(u, np.array([0.0, 1.0, 0.0], np.float32))
    assert np.array_equal(c, np.array([0.07, 0.07, 0.25], np.float32))
    u = uniforms((64, 64, 64), night=1.0, sun=(0.2, -1.0, 0.1))
    c = oracle.sky_color(u, np.array([0.0, 1.0, 0.0], np.float32))
    assert np.array_equal(c, np.array([0.005, 0.006, 0.02], np.float32))
    # looking into the sun (high in the sky, away from beam and towers)
    u = uniforms((64, 64, 64), night=0.0, sun=(0.0, 1.0, 0.0))
    base = np.array([0.07, 0.07, 0.25], np.float32)
    c = oracle.sky_color(u, u[4:7].copy())
    assert np.allclose(c, base + u[8:11] * np.float32(40.35), rtol=1e-6)


# ---- the image ---------------------------------------------------------------
def _render(ow, u, li, rays):
    return ow.render(u, li, rays, threads=4)


def test_debug_views_follow_the_hits():
    s = gallery()
    ow = oracle.World.from_scene(s)
    rays = panorama(64, 32)
    u = uniforms(s.dims)
    li = lights()
    hits = [ow.march_ex(u, r["ro"], r["rd"]) for r in rays]
    normals = {0: (1, 0, 0), 1: (-1, 0, 0), 2: (0, 1, 0), 3: (0, -1, 0), 4: (0, 0, 1), 5: (0, 0, -1), 6: (0, 0, 0)}
    views = {k: _render(ow, uniforms(s.dims, debug=float(k)), li, rays) for k in (1, 3, 4)}
    for k, (h, _, _, _, _) in enumerate(hits):
        if h["face"] == 0xFF:
            for v in views.values():
                assert np.array_equal(v[k, :3], np.array([0.2, 0.0, 0.3], np.float32))
            continue
        n = np.asarray(normals[int(h["face"])], np.float32)
        assert np.array_equal(views[1][k, :3], n * np.float32(0.5) + np.float32(0.5))
        q = np.float32(h["t"]) / np.float32(100.0)
        assert views[3][k, 0] == q - np.floor(q)
        idf = np.float32(h["mat"])
        assert views[4][k, 0] == idf / np.float32(64.0) and views[4][k, 2] == np.float32(0.5)
    assert all(v[:, 3].min() == 1.0 for v in views.values())


def test_emissive_is_albedo_times_emission():
    s = gallery()
    ow = oracle.World.from_scene(s)
    rays = panorama(96, 48)
    li = lights()
    a = _render(ow, uniforms(s.dims), li, rays)
    b = _render(ow, uniforms(s.dims, night=0.9, sun_w=0.0, ao=0.0), lights([]), rays)
    alb = _render(ow, uniforms(s.dims, debug=2.0), li, rays)
    em = 0
    for k, r in enumerate(rays):
        h, trans, glow, _, _ = ow.march_ex(uniforms(s.dims), r["ro"], r["rd"])
        if h["face"] != 0xFF and h["mat"] == 40 and np.all(trans == 1.0):
            em += 1
            assert np.array_equal(a[k, :3], alb[k, :3] * np.float32(6.0))
            assert np.array_equal(b[k, :3], a[k, :3])
    assert em > 5


def test_shadow_equals_sun_off_and_far_light_is_nothing():
    """Floor pixels in the overhang's shadow are bit-identical to the same
    pixels with the sun switched off; pixels in the sun are not."""
    s = gallery()
    ow = oracle.World.from_scene(s)
    rays = panorama(128, 64)
    li = lights([])
    u_sun = uniforms(s.dims, sun=(0.05, 1.0, 0.02), reflections=0.0)
    u_off = u_sun.copy()
    u_off[7] = 0.0
    a = _render(ow, u_sun, li, rays)
    b = _render(ow, u_off, li, rays)
    shadowed = lit = 0
    for k, r in enumerate(rays):
        h, _, _, _, _ = ow.march_ex(u_sun, r["ro"], r["rd"])
        if h["face"] != 2 or h["mat"] != 1:
            continue
        P = r["ro"] + r["rd"] * h["t"]
        sh, _, _, _, _ = ow.march_ex(u_sun, P + np.array([0, 0.5, 0], np.float32), u_sun[4:7], shadow=True)
        if sh["face"] != 0xFF and sh["mat"] == 3:
            shadowed += 1
            assert np.array_equal(f32bits(a[k]), f32bits(b[k]))
        elif sh["face"] == 0xFF:
            lit += 1
            assert np.all(a[k, :3] > b[k, :3])
    assert shadowed > 10 and lit > 10
    # a light whose radius ends before any surface contributes nothing
    far = lights([((48.3, 14.6, 47.7), 0.5, (5.0, 5.0, 5.0), 1.0)])
    c = _render(ow, u_sun, far, rays)
    assert np.array_equal(f32bits(c), f32bits(a))


def test_fine_ao_beside_a_wall():
    """A floor voxel at x = 9 with a wall at x = 10 (all z): its top face has
    su1 = c10 = c11 = 1, so ao = mix(1, 1/3, fu) and the view is
    0.35 + 0.65 (1 - 2/3 fu) with fu = P.x - 9 (coarse AO off)."""
    from pixels_amd.scene import Scene
    vox = np.zeros((32, 32, 32), np.uint16)
    vox[:, :4, :] = 1
    vox[:, 4:20, 10] = 1
    s = Scene.from_dense(vox)
    ow = oracle.World.from_scene(s)
    u = uniforms(s.dims, debug=6.0, ao=0.0)
    xs = np.linspace(9.05, 9.95, 19)
    rays = np.zeros(xs.size, oracle.RAY_DTYPE)
    rays["ro"] = np.stack([xs, np.full_like(xs, 12.0), np.full_like(xs, 16.3)], -1).astype(np.float32)
    rays["rd"] = np.array([0.0, -1.0, 0.0], np.float32)
    rays["max_t"] = np.float32(1e9)
    img = _render(ow, u, lights([]), rays)
    for k, x in enumerate(rays["ro"][:, 0]):
        fu = float(np.float32(x) - np.float32(9.0))
        exp = 0.35 + 0.65 * (1.0 - (2.0 / 3.0) * fu)
        assert abs(float(img[k, 0]) - exp) < 2e-6, (x, img[k, 0], exp)


def test_glass_pane_transmittance_and_glow():
    """Head-on through a one-voxel pane: colour = T * (colour without the
    pane) + glow, T = (1 - F)(0.4 + 0.6 tint) g with F = 0.04, glow = F *
    sky(reflected) * 0.35 + glint (the floor behind is shaded identically)."""
    from pixels_amd.scene import Scene
    m = palette()
    vox = np.zeros((32, 32, 32), np.uint16)
    vox[:, :4, :] = 1
    vox2 = vox.copy()
    vox2[:, 4:30, 20] = 43
    s1, s2 = Scene.from_dense(vox, m), Scene.from_dense(vox2, m)
    o1, o2 = oracle.World.from_scene(s1), oracle.World.from_scene(s2)
    u = uniforms(s1.dims, sun=(-0.3, 0.8, 0.2), reflections=0.0)
    rays = np.zeros(8, oracle.RAY_DTYPE)
    d = np.array([1.0, -0.02, 0.0])
    rays["ro"] = np.stack([np.full(8, 5.0), np.full(8, 10.0), np.linspace(8, 24, 8)], -1).astype(np.float32)
    rays["rd"] = (d / np.linalg.norm(d)).astype(np.float32)
    rays["max_t"] = np.float32(1e9)
    a = _render(o1, u, lights([]), rays)
    b = _render(o2, u, lights([]), rays)
    tint = m.reshape(256, 20)[43, 0:3].astype(np.float64)
    g = float(m.reshape(256, 20)[43, 13])
    rd = rays["rd"][0].astype(np.float64)
    F = 0.04 + 0.96 * (1.0 - abs(rd[0])) ** 5
    T = (1.0 - F) * (0.4 + 0.6 * tint) * g
    r = rd.copy()
    r[0] = -r[0]
    sky = oracle.sky_color(u, r.astype(np.float32)).astype(np.float64)
    glint = max(float(np.dot(r, u[4:7])), 0.0) ** 200.0 * u[8:11] * u[7] * 2.0
    glow = F * sky * 0.35 + glint * 1.0
    for k in range(8):
        exp = T * a[k, :3] + glow
        assert np.allclose(b[k, :3], exp, rtol=2e-5, atol=1e-6), (k, b[k], exp)
