#!/usr/bin/env python3
"""Benchmark of the brick-map hot path on one MI355X.

A step is one frame of the renderer's ray work on the full-scale scene
(1280x352x960 voxels, 2000 seeded primitives, structural floor/ceiling
slabs, 10% glass): 1920x1080 primary rays through march() plus, per
primary hit, the sun shadow ray and shadow rays to 8 point lights (shade(),
raymarch.wgsl:323-356).  Inputs (rays, world) are resident in HBM before
timing starts.  Prints one JSON line on stdout; progress goes to stderr.

value = (primary + shadow rays traced) / frame time, in Grays/s.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Grays/s primary+shadow @1280x352x960 bit-exact; island-pass ms; % HBM roofline"
HBM_PEAK_GBS = 8000.0
BYTES_IN, BYTES_OUT, BYTES_STEP = 28, 13, 4
COPY_BYTES = 256 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the parity check")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the hot path is per-GPU work (rays shard with the brick map "
              "replicated); run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    from pixels_amd import GpuWorld, Scene, Simulation
    from pixels_amd import _native as N
    from pixels_amd.scene import LARGE

    t0 = time.time()
    scene = Scene.build(LARGE, seed=args.seed)
    W, H = args.width, args.height
    rays = scene.camera_rays(W, H, 90.0, seed=args.seed, tile=8)
    u, lights = scene.frame_inputs(W, H, seed=args.seed)
    log("scene %s fill %.3f pool bricks %d built in %.1fs" % (scene.dims, scene.fill, scene.pool_used,
                                                              time.time() - t0))
    gw = GpuWorld(scene)
    sim = Simulation(gw)
    gw.set_uniforms(u)
    gw.set_lights(lights)
    slots = sim.frame_slots()
    n = W * H
    dev = torch.device("cuda:0")
    d_rays = torch.from_numpy(rays.view(np.uint8).copy()).to(dev)
    d_prim = torch.empty(n * 16, dtype=torch.uint8, device=dev)
    d_sh = torch.empty(n * slots * 16, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(gw.stream)

    for _ in range(args.warmup):
        sim.trace_frame_device(d_rays, n, d_prim, d_sh)
    gw.sync()
    log("warmup done")

    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    gw.sync()
    gw.profile_begin()
    ev0.record(stream)
    for _ in range(args.steps):
        sim.trace_frame_device(d_rays, n, d_prim, d_sh)
    ev1.record(stream)
    ev1.synchronize()
    gw.sync()
    torch.cuda.synchronize()
    prof = gw.profile_end()
    frame_ms = ev0.elapsed_time(ev1) / args.steps
    log("timed %d frames: %.3f ms/frame" % (args.steps, frame_ms))
    # a plain device copy of COPY_BYTES: under rocprofv3 --pmc its FETCH_SIZE /
    # WRITE_SIZE give the scale of those counters against known bytes
    cal_src = torch.zeros(COPY_BYTES, dtype=torch.uint8, device=dev)
    cal_dst = torch.empty_like(cal_src)
    cal_dst.copy_(cal_src)
    torch.cuda.synchronize()
    del cal_src, cal_dst

    prim = d_prim.cpu().numpy().view(N.HIT_DTYPE)
    sh = d_sh.cpu().numpy().view(N.HIT_DTYPE)
    cast = sh["face"] != N.FACE_NOT_CAST
    n_cast = int(cast.sum())
    total = n + n_cast
    value = total / (frame_ms * 1e-3) / 1e9

    steps_prim = int(prim["steps"].astype(np.int64).sum())
    steps_sh = int(sh["steps"][cast].astype(np.int64).sum())
    bytes_prim = n * (BYTES_IN + BYTES_OUT) + BYTES_STEP * steps_prim
    bytes_sh = n_cast * (BYTES_IN + BYTES_OUT) + BYTES_STEP * steps_sh
    k_ms = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in prof.items()}
    ms_p = k_ms.get(N.PROF_PRIMARY, 0.0)
    ms_s = k_ms.get(N.PROF_SHADOW, 0.0)
    names = ("k_trace_shadow_queue", "k_frame_primary")
    if ms_s >= ms_p:
        dom, dom_bytes, dom_ms = names[0], bytes_sh, ms_s
    else:
        dom, dom_bytes, dom_ms = names[1], bytes_prim, ms_p
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                "kernel_ms": round(dom_ms, 4), "algorithmic_bytes": dom_bytes,
                "kernels_ms": {"primary": round(ms_p, 4), "shadow": round(ms_s, 4)}}

    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        threads = args.cpu_threads or min(os.cpu_count() or 1, 64)
        ow = oracle.World.from_scene(scene)
        log("cpu baseline: oracle frame on %d threads" % threads)
        t1 = time.time()
        rp, rs, rnc = ow.frame(u, lights, rays, threads=threads)
        cpu_s = time.time() - t1
        cpu = {"value": round((n + rnc) / cpu_s / 1e9, 6), "unit": "Grays/s", "cores": threads, "kind": "port",
               "sample": "the whole %dx%d frame (%d primary + %d shadow rays), one pass, %.2f s wall" %
                         (W, H, n, rnc, cpu_s)}

        def same(a, b):
            return bool(np.array_equal(a.view(np.uint8), b.view(np.uint8)))
        parity = same(prim, rp) and same(sh, rs) and rnc == n_cast
        log("bit-exact vs oracle: %s" % parity)

    gw.close()
    island = island_leg(args)

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(frame_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded scene and camera, no checkpoint or dataset",
        "config": {"workload": "frame: 1920x1080 primary + sun + 8 point-light shadow rays, 1280x352x960 grid",
                   "grid": list(scene.dims), "primitives": scene.n_prims, "pool_bricks": scene.pool_used,
                   "fill": round(scene.fill, 4), "primary_rays": n, "shadow_rays": n_cast,
                   "bit_exact_vs_oracle": parity},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "island_pass": island,
    }
    print(json.dumps(out), flush=True)
    return 0


def island_leg(args):
    """Island pass (carve + pane flood + maint + labelling) for 64 impacts on
    the 512x256x512 scene with a 16-voxel structural floor; the world is
    restored from a device-side snapshot before every timed pass."""
    from pixels_amd import GpuWorld, Scene, Simulation
    from pixels_amd.scene import SMALL_FLOOR
    from pixels_amd.world import impacts_from_hits

    scene = Scene.build(SMALL_FLOOR, seed=3)
    gw = GpuWorld(scene)
    sim = Simulation(gw)
    rng = np.random.default_rng(8)
    rays = scene.camera_rays(640, 360, 90.0, seed=3)
    pick = rays[rng.permutation(rays.shape[0])[:400]]
    hits = sim.trace(0, pick)
    ok = hits["face"] != 0xFF
    pick, hits = pick[ok][:64], hits[ok][:64]
    r = rng.uniform(8, 96, pick.shape[0]).astype(np.float32)
    impacts = impacts_from_hits(scene.dims, pick, hits, r)
    gw.snapshot()
    sim.push_impacts(impacts)
    sim.island_pass()
    gw.sync()
    reps = []
    for _ in range(max(3, args.steps // 4)):
        gw.restore()
        sim.push_impacts(impacts)
        gw.sync()
        gw.profile_begin()
        sim.island_pass()
        prof = gw.profile_end()
        reps.append(prof[3][0])
    keys, cnt, n = sim.island_labels()
    ctrl, _, _ = sim.read_ctrl()
    from pixels_amd import CTRL
    res = {"ms": round(float(np.median(reps)), 4), "ms_min": round(float(np.min(reps)), 4), "reps": len(reps),
           "impacts": int(impacts.shape[0]), "grid": list(scene.dims), "unsupported_components": n,
           "destroyed": int(ctrl[CTRL["S_DESTROYED"]]) - 0, "target_ms": 2.0}
    log("island pass: %.3f ms median over %d passes, %d unsupported components" % (res["ms"], len(reps), n))
    if not args.no_cpu:
        import oracle
        dense = scene.dense()
        kind = np.where(scene.grid == 0, 0, np.where(scene.grid & 0x80000000, 1, 2)).astype(np.uint8)
        t1 = time.time()
        ovox, okind, okeys, ocnt, on, stats = oracle.island_pass(scene.dims, dense, kind, scene.mats, impacts)
        res["cpu_oracle_s"] = round(time.time() - t1, 3)
        vox = sim.probe((0, 0, 0), scene.dims).reshape(ovox.shape)
        res["bit_exact_vs_oracle"] = bool(n == on and np.array_equal(keys, okeys) and np.array_equal(cnt, ocnt)
                                          and np.array_equal(vox, ovox.astype(np.uint32)))
        log("island pass bit-exact vs oracle: %s" % res["bit_exact_vs_oracle"])
    gw.close()
    return res


if __name__ == "__main__":
    sys.exit(main())
