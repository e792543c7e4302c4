"""GPU parity of the shading path (px_render / px_render_frame_device,
fsMain raymarch.wgsl:387-450) with the oracle's restatement (oracle/px_shade.c),
bit for bit on every RGBA float:

* the gallery world (every material pattern, shade bytes, emission, a mirror
  and a glossy reflector, two glass panes, an overhang) seen as a full
  panorama, by day and by night, reflections and lamp shadows on and off;
* every debug view 1-9;
* the seeded 512x256x512 scene with its frame inputs and reflections on;
* the AO density texture against the oracle's, before and after an island
  pass, and the render of the edited world;
* the device entry point against the host one;
* ragged pixel counts (1, 63, 65, 257), no lights and the 32-light maximum,
  an empty world (all sky) and zero rays."""
import numpy as np
import pytest

import oracle
from tests.test_shade_oracle import f32bits, gallery, lights, panorama, uniforms

pytestmark = pytest.mark.gpu


def _open(scene):
    from pixels_amd import GpuWorld, Simulation
    gw = GpuWorld(scene)
    return gw, Simulation(gw)


def _compare(got, exp, what):
    g, e = f32bits(got), f32bits(exp)
    bad = np.nonzero((g != e).any(axis=1))[0]
    assert bad.size == 0, "%s: %d of %d pixels differ, first %s: gpu %r oracle %r" % (
        what, bad.size, g.shape[0], bad[:5], got[bad[:3]], exp[bad[:3]])


@pytest.fixture(scope="module")
def gal():
    s = gallery()
    gw, sim = _open(s)
    yield s, gw, sim, oracle.World.from_scene(s)
    gw.close()


VARIANTS = {
    "day": dict(night=0.0),
    "dusk": dict(night=0.3, time=12.5),
    "night_no_refl": dict(night=1.0, reflections=0.0, lamp_shadows=0.0, sun=(0.3, -0.6, 0.2)),
    "low_sun": dict(night=0.1, sun=(0.9, 0.12, -0.4), ao=1.0, time=3.25),
}


@pytest.mark.parametrize("name", sorted(VARIANTS))
def test_gallery_matches_oracle(gal, name):
    s, gw, sim, ow = gal
    u = uniforms(s.dims, **VARIANTS[name])
    li = lights()
    rays = panorama(256, 128)
    print("case: gallery %s, %d rays" % (name, rays.shape[0]), flush=True)
    gw.set_uniforms(u)
    gw.set_lights(li)
    got = sim.render(rays)
    exp = ow.render(u, li, rays)
    _compare(got, exp, name)
    assert np.isfinite(got).all() and len(np.unique(f32bits(got[:, 0]))) > 1000


@pytest.mark.parametrize("dbg", range(1, 10))
def test_debug_views_match_oracle(gal, dbg):
    s, gw, sim, ow = gal
    u = uniforms(s.dims, debug=float(dbg))
    li = lights()
    rays = panorama(128, 64)
    print("case: debug view %d" % dbg, flush=True)
    gw.set_uniforms(u)
    gw.set_lights(li)
    _compare(sim.render(rays), ow.render(u, li, rays), "debug %d" % dbg)


def test_seeded_scene_matches_oracle(small_scene):
    s = small_scene
    gw, sim = _open(s)
    try:
        W, H = 320, 180
        rays = s.camera_rays(W, H, 90.0, seed=4)
        u, li = s.frame_inputs(W, H, seed=4)
        u[20] = 1.0
        print("case: seeded scene %s, %dx%d" % (s.dims, W, H), flush=True)
        gw.set_uniforms(u)
        gw.set_lights(li)
        ow = oracle.World.from_scene(s)
        assert np.array_equal(sim.read_density(), ow.density())
        _compare(sim.render(rays), ow.render(u, li, rays), "seeded scene")
    finally:
        gw.close()


def test_density_and_render_after_edits(floor_scene):
    """The island pass changes the world; the density texture and the image
    follow the edited voxels."""
    from pixels_amd.scene import Scene
    from pixels_amd.world import impacts_from_hits
    s = floor_scene
    gw, sim = _open(s)
    try:
        rng = np.random.default_rng(8)
        rays = s.camera_rays(640, 360, 90.0, seed=3)
        pick = rays[rng.permutation(rays.shape[0])[:400]]
        hits = sim.trace(0, pick)
        ok = hits["face"] != 0xFF
        pick, hits = pick[ok][:64], hits[ok][:64]
        r = rng.uniform(8, 96, pick.shape[0]).astype(np.float32)
        sim.push_impacts(impacts_from_hits(s.dims, pick, hits, r))
        d0 = sim.read_density()
        print("case: floor scene island pass, then density + render", flush=True)
        sim.island_pass()
        vox = sim.probe((0, 0, 0), s.dims).reshape(s.dims[2], s.dims[1], s.dims[0]).astype(np.uint16)
        ow = oracle.World.from_scene(Scene.from_dense(vox, s.mats))
        d1 = sim.read_density()
        assert np.array_equal(d1, ow.density()) and not np.array_equal(d0, d1)
        W, H = 256, 144
        cam_rays = s.camera_rays(W, H, 90.0, seed=3)
        u, li = s.frame_inputs(W, H, seed=3)
        u[20] = 1.0
        gw.set_uniforms(u)
        gw.set_lights(li)
        _compare(sim.render(cam_rays), ow.render(u, li, cam_rays), "edited floor scene")
    finally:
        gw.close()


def test_device_entry_equals_host(gal):
    import torch
    s, gw, sim, ow = gal
    u = uniforms(s.dims)
    li = lights()
    gw.set_uniforms(u)
    gw.set_lights(li)
    rays = panorama(96, 48)
    host = sim.render(rays)
    dev = torch.device("cuda:0")
    d_rays = torch.from_numpy(rays.view(np.uint8).copy()).to(dev)
    d_rgba = torch.zeros(rays.shape[0] * 4, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    sim.render_device(d_rays, rays.shape[0], d_rgba)
    gw.sync()
    got = d_rgba.cpu().numpy().reshape(-1, 4)
    assert np.array_equal(f32bits(got), f32bits(host))


@pytest.mark.parametrize("n", [1, 63, 65, 257])
def test_ragged_counts_match_oracle(gal, n):
    """Pixel counts that do not fill a wavefront or a workgroup."""
    s, gw, sim, ow = gal
    u = uniforms(s.dims)
    li = lights()
    rays = panorama(64, 32)[np.random.default_rng(n).permutation(64 * 32)[:n]]
    print("case: ragged render n=%d" % n, flush=True)
    gw.set_uniforms(u)
    gw.set_lights(li)
    _compare(sim.render(rays), ow.render(u, li, rays), "n=%d" % n)


@pytest.mark.parametrize("n_lights", [0, 32])
def test_light_counts_match_oracle(gal, n_lights):
    s, gw, sim, ow = gal
    rng = np.random.default_rng(n_lights)
    entries = [(tuple(rng.uniform(4, 90, 3)), float(rng.uniform(10, 80)), tuple(rng.uniform(0.2, 1.0, 3)),
                float(rng.integers(0, 2))) for _ in range(n_lights)]
    u = uniforms(s.dims)
    li = lights(entries)
    rays = panorama(96, 48)
    print("case: render with %d lights" % n_lights, flush=True)
    gw.set_uniforms(u)
    gw.set_lights(li)
    _compare(sim.render(rays), ow.render(u, li, rays), "%d lights" % n_lights)


def test_empty_world_is_sky_and_zero_rays():
    from pixels_amd.scene import Scene
    from tests.test_shade_oracle import palette
    s = Scene.from_dense(np.zeros((32, 32, 32), np.uint16), palette())
    gw, sim = _open(s)
    try:
        u = uniforms(s.dims, night=0.6)
        gw.set_uniforms(u)
        gw.set_lights(lights())
        rays = panorama(64, 32, cam=(16.0, 16.0, 16.0))
        print("case: empty world", flush=True)
        got = sim.render(rays)
        ow = oracle.World.from_scene(s)
        _compare(got, ow.render(u, lights(), rays), "empty world")
        assert not sim.read_density().any()
        assert sim.render(rays[:0]).shape == (0, 4)
    finally:
        gw.close()
