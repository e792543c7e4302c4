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
          "cls: [typeof m.GpuWorld, typeof m.Simulation], ctrl: m.CTRL.S_DESTROYED}));"
          % os.path.join(ROOT, "pixels_amd", "node", "index.js"))
    r = subprocess.run([NODE, "-e", js], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["fns"] == sorted(["create", "destroy", "setUniforms", "setLights", "trace", "probe", "readCtrl",
                                 "pushImpacts", "islandPass", "islandLabels", "lastError"])
    assert out["cls"] == ["function", "function"]
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
  console.log(JSON.stringify({ n: lab.keys.length, keys: Array.from(lab.keys), counts: Array.from(lab.counts), destroyed: st.destroyed }));
  gw.destroy();
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
