"""The N-API addon loads in Node and exposes the GpuWorld / Simulation mirror
(no compute without a GPU); on a GPU box a Node script traces rays and runs
an island pass, and its results are checked against the oracle."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "pixels_amd", "node", "pixels_napi.node")

needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the addon is missing")


@needs_node
def test_addon_loads_and_exports():
    js = ("const m = require(%r); console.log(JSON.stringify({fns: Object.keys(m.addon).sort(), "
          "cls: [typeof m.GpuWorld, typeof m.Simulation, typeof m.Renderer], ctrl: m.CTRL.S_DESTROYED}));"
          % os.path.join(ROOT, "pixels_amd", "node", "index.js"))
    r = subprocess.run([NODE, "-e", js], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["fns"] == sorted(["create", "destroy", "setUniforms", "setLights", "trace", "probe", "readCtrl",
                                 "pushImpacts", "islandPass", "islandLabels", "lastError", "createFromPrims",
                                 "simStep", "particlesInit", "spawn", "readParticles", "readAlive", "pushWrites",
                                 "readChunks", "render", "readDensity"])
    assert out["cls"] == ["function", "function", "function"]
    assert out["ctrl"] == 25


@needs_node
def test_addon_error_without_device_or_bad_args():
    js = ("const m = require(%r); try { m.addon.create({dims:[8,8,8], grid:new Uint32Array(1), "
          "mats:new Float32Array(16)}); console.log('no-throw'); } catch (e) { console.log('threw:' + e.message); }"
          % os.path.join(ROOT, "pixels_amd", "node", "index.js"))
    r = subprocess.run([NODE, "-e", js], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().startswith("threw:")


@needs_node
def test_addon_rejects_short_buffers():
    """create() checks the grid and pool lengths against dims and poolUsed
    before the library reads them (no device needed: it throws first)."""
    js = ("const m = require(%r); const mats = new Float32Array(256 * 20); const out = [];"
          "try { m.addon.create({dims:[16,16,16], grid:new Uint32Array(7), mats}); out.push('no-throw'); }"
          " catch (e) { out.push(e.message); }"
          "try { m.addon.create({dims:[16,16,16], grid:new Uint32Array(8), pool:new Uint16Array(600), poolUsed:2,"
          " mats}); out.push('no-throw'); } catch (e) { out.push(e.message); }"
          "console.log(JSON.stringify(out));" % os.path.join(ROOT, "pixels_amd", "node", "index.js"))
    r = subprocess.run([NODE, "-e", js], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out[0].startswith("grid:") and out[1].startswith("pool:"), out


@needs_node
@pytest.mark.gpu
def test_node_host_traces_and_edits(tmp_path):
    import oracle
    from pixels_amd import Scene
    from pixels_amd.scene import SMALL_FLOOR
    s = Scene.build(SMALL_FLOOR, seed=5, dims=(128, 64, 128), n_prims=8)
    rays = s.camera_rays(96, 64, 90.0, seed=5)
    imp = np.zeros((2, 16), np.float32)
    imp[0, :5] = (40.0, 20.0, 40.0, 12.0, 12.0)
    imp[1, :5] = (90.0, 30.0, 70.0, 20.0, 20.0)
    for name, arr in (("grid", s.grid), ("pool", s.pool), ("mats", s.mats), ("rays", rays), ("imp", imp)):
        np.ascontiguousarray(arr).tofile(str(tmp_path / name))
    js = r"""
const fs = require('fs');
const m = require(%r);
const d = %r;
const rd = (n, T) => { const b = fs.readFileSync(d + '/' + n); return new T(b.buffer, b.byteOffset, b.byteLength / T.BYTES_PER_ELEMENT); };
const world = { dx: 128, dy: 64, dz: 128, grid: rd('grid', Uint32Array), pool: rd('pool', Uint16Array), poolUsed: %d };
const gw = new m.GpuWorld(0, world, rd('mats', Float32Array));
const sim = new m.Simulation(gw);
const hits = sim.trace(m.RAY.PRIMARY, rd('rays', Uint8Array));
fs.writeFileSync(d + '/hits', Buffer.from(hits.buffer));
sim.pushImpacts(rd('imp', Float32Array));
sim.encode();
const lab = sim.islandLabels();
const st = sim.afterSubmit();
sim.probe([0, 0, 0], [128, 64, 128]).then((vox) => {
  fs.writeFileSync(d + '/vox', Buffer.from(vox.buffer));
  const raw = gw.ctx;
  gw.destroy();
  let after = 'no-throw';  // a call on a destroyed context throws, it does not touch freed memory
  try { m.addon.trace(raw, 0, rd('rays', Uint8Array)); } catch (e) { after = e.message; }
  console.log(JSON.stringify({ n: lab.keys.length, keys: Array.from(lab.keys), counts: Array.from(lab.counts), destroyed: st.destroyed, after: after }));
});
""" % (os.path.join(ROOT, "pixels_amd", "node", "index.js"), str(tmp_path), s.pool_used)
    print("case: node host", flush=True)
    r = subprocess.run([NODE, "-e", js], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    hits = np.fromfile(str(tmp_path / "hits"), oracle.HIT_DTYPE)
    ref = oracle.World.from_scene(s).trace(0, rays)
    assert np.array_equal(hits.view(np.uint8), ref.view(np.uint8))
    kind = np.where(s.grid == 0, 0, np.where(s.grid & 0x80000000, 1, 2)).astype(np.uint8)
    ovox, _, okeys, ocnt, on, stats = oracle.island_pass(s.dims, s.dense(), kind, s.mats, imp)
    vox = np.fromfile(str(tmp_path / "vox"), np.uint32).reshape(ovox.shape)
    assert np.array_equal(vox, ovox.astype(np.uint32))
    assert out["n"] == on and out["keys"] == okeys.tolist() and out["counts"] == ocnt.tolist()
    assert out["destroyed"] == int(stats[0])
    assert "destroyed" in out["after"]


@needs_node
@pytest.mark.gpu
def test_node_host_scene_build_render_and_debris(tmp_path):
    """The Node host builds a world from primitives on the GPU, renders it
    (fsMain colour), and runs a debris frame through the stage mask; each
    result equals the oracle's."""
    import oracle
    from pixels_amd import _native as N
    from pixels_amd.scene import LARGE, Scene, debris
    from tests.test_shade_oracle import lights, uniforms
    s = Scene.build(LARGE, seed=4, dims=(128, 64, 96), n_prims=40)
    prims = s.prims()
    rays = s.camera_rays(64, 48, 90.0, seed=2)
    u = uniforms(s.dims)
    u[24] = np.float32(1.0 / 60.0)
    li = lights([((40.0, 30.0, 40.0), 60.0, (1.0, 0.8, 0.6), 1.0)])
    recs = debris(s.dims, 500, seed=3, speed=80.0)
    for name, arr in (("prims", prims), ("mats", s.mats), ("rays", rays), ("u", u), ("li", li), ("recs", recs)):
        np.ascontiguousarray(arr).tofile(str(tmp_path / name))
    js = r"""
const fs = require('fs');
const m = require(%r);
const d = %r;
const rd = (n, T) => { const b = fs.readFileSync(d + '/' + n); return new T(b.buffer, b.byteOffset, b.byteLength / T.BYTES_PER_ELEMENT); };
const gw = m.GpuWorld.fromPrims(0, [128, 64, 96], rd('prims', Uint8Array), rd('mats', Float32Array));
gw.writeUniforms(rd('u', Float32Array));
gw.writeLights(rd('li', Float32Array));
const img = new m.Renderer(gw).render(rd('rays', Uint8Array));
fs.writeFileSync(d + '/img', Buffer.from(img.buffer));
const sim = new m.Simulation(gw);
sim.particlesInit(4096);
sim.encode(m.STAGE.FRAME_BEGIN);
sim.spawn(rd('recs', Uint8Array));
sim.encode(m.STAGE.PARTICLES);
fs.writeFileSync(d + '/parts', Buffer.from(sim.readParticles(0, 4096).buffer));
fs.writeFileSync(d + '/alive', Buffer.from(sim.readAlive().buffer));
sim.probe([0, 0, 0], [128, 64, 96]).then((vox) => {
  fs.writeFileSync(d + '/vox', Buffer.from(vox.buffer));
  gw.destroy();
  console.log('ok');
});
""" % (os.path.join(ROOT, "pixels_amd", "node", "index.js"), str(tmp_path))
    print("case: node host scene build + render + debris", flush=True)
    r = subprocess.run([NODE, "-e", js], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    vox, _ = oracle.raster(s.dims, prims)
    got = np.fromfile(str(tmp_path / "vox"), np.uint32).reshape(vox.shape)
    assert np.array_equal(got, vox.astype(np.uint32))
    ow = oracle.World.from_scene(Scene.from_dense(vox, s.mats))
    img = np.fromfile(str(tmp_path / "img"), np.float32).reshape(-1, 4)
    assert np.array_equal(img.view(np.uint32), ow.render(u, li, rays).view(np.uint32))
    P = oracle.Particles(4096)
    P.frame_begin()
    P.spawn(recs)
    P.step(ow, np.float32(1.0 / 60.0), np.float32(u[25]), np.float32(u[3]))
    parts = np.fromfile(str(tmp_path / "parts"), N.PARTICLE_DTYPE)
    assert parts.tobytes() == P.parts.tobytes()
    assert np.array_equal(np.fromfile(str(tmp_path / "alive"), np.uint32), P.alive_list())
