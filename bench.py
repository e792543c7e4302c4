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
# windows of the learned sort in the batched comparisons: 2^18 consecutive
# queries, each sorted by (direction face, predicted cost) on its own
WINDOW_LOG2 = 18


TRAFFIC_FILE = "profiles/hbm_traffic_frame_1920x1080.md"
DEBRIS_TRAFFIC_FILE = "profiles/debris_step.md"


def profiled_traffic(kernel, source=TRAFFIC_FILE):
    """HBM bytes per launch and L2 hit % of `kernel` from a committed
    rocprofv3 --pmc summary (FETCH_SIZE / WRITE_SIZE scaled by a 256 MiB copy);
    (None, None) if the summary has no line for it."""
    path = os.path.join(ROOT, source)
    traffic = hit = None
    if os.path.exists(path):
        for line in open(path):
            f = line.split()
            if len(f) == 3 and f[1] == kernel:
                if f[0] == "traffic_bytes":
                    traffic = int(f[2])
                elif f[0] == "l2_hit_pct":
                    hit = float(f[2])
    return traffic, hit


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
    ap.add_argument("--no-batch", action="store_true", help="skip the batched mixed-query leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned ray-cost binning leg")
    ap.add_argument("--no-debris", action="store_true", help="skip the debris-particle leg")
    ap.add_argument("--no-scene", action="store_true", help="skip the GPU scene-build leg")
    ap.add_argument("--no-render", action="store_true", help="skip the shading-to-an-image leg")
    ap.add_argument("--layout", choices=("planar", "pixel"), default="planar",
                    help="shadow outputs of the frame: slot-planar (s * n + i) or pixel-major (i * slots + s)")
    ap.add_argument("--compare-layout", action="store_true", help="also time the frame with the other layout")
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
    planar = args.layout == "planar"
    sim.set_frame_layout(planar)

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
    if planar:  # to pixel-major (slot s of ray i at i * slots + s), as the oracle writes it
        sh = np.ascontiguousarray(sh.reshape(slots, n).T).reshape(-1)
    other_ms = None
    if args.compare_layout:
        # the other layout, timed the same way (same records)
        sim.set_frame_layout(not planar)
        for _ in range(args.warmup):
            sim.trace_frame_device(d_rays, n, d_prim, d_sh)
        gw.sync()
        ev0.record(stream)
        for _ in range(args.steps):
            sim.trace_frame_device(d_rays, n, d_prim, d_sh)
        ev1.record(stream)
        ev1.synchronize()
        other_ms = round(ev0.elapsed_time(ev1) / args.steps, 4)
        sim.set_frame_layout(planar)
        log("other layout: %.3f ms/frame" % other_ms)
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
    traffic, l2_hit = profiled_traffic(dom)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "l2_hit_pct": l2_hit,
                "traffic_source": TRAFFIC_FILE if traffic is not None else None, "kernel": dom,
                "kernel_ms": round(dom_ms, 4), "algorithmic_bytes": dom_bytes,
                "kernels_ms": {"primary": round(ms_p, 4), "shadow": round(ms_s, 4)}}

    cpu = None
    parity = None
    if not args.no_cpu:
        import oracle
        threads = args.cpu_threads or min(os.cpu_count() or 1, 64)
        ow = oracle.World.from_scene(scene)
        log("cpu baseline: oracle frame on %d threads, 1 warm-up + 5 timed passes" % threads)
        rp, rs, rnc = ow.frame(u, lights, rays, threads=threads)  # warm-up (and the parity result)
        walls = []
        for _ in range(5):
            t1 = time.time()
            ow.frame(u, lights, rays, threads=threads)
            walls.append(time.time() - t1)
        cpu_s = float(np.median(walls))
        cpu = {"value": round((n + rnc) / cpu_s / 1e9, 6), "unit": "Grays/s", "cores": threads, "kind": "port",
               "sample": "the whole %dx%d frame (%d primary + %d shadow rays): median of 5 passes after 1 warm-up, "
                         "%.3f s wall each (range %.3f-%.3f); oracle built -O3 -ffp-contract=off" %
                         (W, H, n, rnc, cpu_s, min(walls), max(walls))}

        def same(a, b):
            return bool(np.array_equal(a.view(np.uint8), b.view(np.uint8)))
        parity = same(prim, rp) and same(sh, rs) and rnc == n_cast
        log("bit-exact vs oracle: %s" % parity)

    debris_res = None if args.no_debris else debris_leg(args, scene, gw, u, rays, prim)
    render_res = None if args.no_render else render_leg(args, scene, gw, u, lights, d_rays, n)
    gw.close()
    scene_build = None if args.no_scene else scene_leg(args, scene)
    island = island_leg(args)
    batched = None if args.no_batch else batch_leg(args)
    learned = None if args.no_learned else learned_leg(args)

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
                   "bit_exact_vs_oracle": parity,
                   "shadow_layout": "slot-planar" if planar else "pixel-major",
                   "other_layout_ms_per_step": other_ms},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "island_pass": island,
        "batched": batched,
        "learned_binning": learned,
        "debris": debris_res,
        "scene_build": scene_build,
        "render": render_res,
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
    # the same pass as one frame stage (PX_STAGE_EDIT), recorded once as a HIP
    # graph and replayed: the ~25 launches without the host between them
    from pixels_amd import _native as N
    gw.restore()
    sim.push_impacts(impacts)
    sim.encode(N.STAGE_EDIT)
    gw.sync()
    graph = sim.capture(N.STAGE_EDIT)
    greps = []
    for _ in range(max(3, args.steps // 4)):
        gw.restore()
        sim.push_impacts(impacts)
        gw.sync()
        gw.profile_begin()
        graph.launch()
        prof = gw.profile_end()
        greps.append(prof[N.PROF_GRAPH][0])
    graph.close()
    gkeys, gcnt, gn = sim.island_labels()
    gctrl, _, _ = sim.read_ctrl()
    from pixels_amd import CTRL
    res = {"ms": round(float(np.median(reps)), 4), "ms_min": round(float(np.min(reps)), 4), "reps": len(reps),
           "ms_graph": round(float(np.median(greps)), 4), "ms_graph_min": round(float(np.min(greps)), 4),
           "graph_equals_eager": bool(gn == n and np.array_equal(gkeys, keys) and np.array_equal(gcnt, cnt)
                                      and np.array_equal(gctrl, ctrl)),
           "impacts": int(impacts.shape[0]), "grid": list(scene.dims), "unsupported_components": n,
           "destroyed": int(ctrl[CTRL["S_DESTROYED"]]) - 0, "target_ms": 2.0}
    # algorithmic bytes of the pass: one occupancy read (4 B per cell + 1 KiB
    # per pool brick) + 8 B per solid voxel, on the world before the pass
    solid = int(np.count_nonzero(scene.dense()))
    alg = 4 * scene.grid.size + 1024 * scene.pool_used + 8 * solid
    ach = alg / (res["ms"] * 1e-3) / 1e9
    res["roofline"] = {"bound": "hbm", "algorithmic_bytes": alg, "solid_voxels": solid,
                       "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(ach / HBM_PEAK_GBS, 4)}
    achg = alg / (res["ms_graph"] * 1e-3) / 1e9
    res["roofline"].update({"achieved_graph": round(achg, 1), "frac_graph": round(achg / HBM_PEAK_GBS, 4)})
    log("island pass: %.3f ms median over %d passes (%.3f ms replayed as a graph, same result: %s), "
        "%d unsupported components" % (res["ms"], len(reps), res["ms_graph"], res["graph_equals_eager"], n))
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
    # extraction and falling chunks continue from the snapshot (after the comparison above)
    res["extraction"] = extraction_leg(args, scene, gw, sim, impacts)
    gw.close()
    return res


def render_leg(args, scene, gw, u, lights, d_rays, n):
    """fsMain to an image on the full-scale scene: the frame's rays with
    shading, glass glow, reflections (U.flags.x = 1) and the sky, RGBA floats
    into HBM; per-kernel times, parity of a pixel sample with the oracle."""
    import torch

    from pixels_amd import Simulation
    from pixels_amd import _native as N
    sim = Simulation(gw)
    ur = u.copy()
    ur[20] = 1.0
    gw.set_uniforms(ur)
    gw.set_lights(lights)
    d_rgba = torch.empty(n * 4, dtype=torch.float32, device=torch.device("cuda:0"))
    torch.cuda.synchronize()
    for _ in range(max(1, args.warmup)):
        sim.render_device(d_rays, n, d_rgba)
    gw.sync()
    steps = max(3, args.steps // 2)
    gw.profile_begin()
    t1 = time.time()
    for _ in range(steps):
        sim.render_device(d_rays, n, d_rgba)
    gw.sync()
    wall = (time.time() - t1) / steps
    prof = gw.profile_end()
    k_ms = {k: (v[0] / steps) for k, v in prof.items() if v[1]}
    ms = sum(k_ms.get(k, 0.0) for k in (N.PROF_RENDER_PRIMARY, N.PROF_RENDER_SHADOW, N.PROF_REFLECT, N.PROF_SHADE))
    res = {"ms": round(ms, 4), "wall_ms": round(wall * 1e3, 4), "mpix_s": round(n / (ms * 1e-3) / 1e6, 1),
           "kernels_ms": {"primary": round(k_ms.get(N.PROF_RENDER_PRIMARY, 0.0), 4),
                          "shadow": round(k_ms.get(N.PROF_RENDER_SHADOW, 0.0), 4),
                          "reflect": round(k_ms.get(N.PROF_REFLECT, 0.0), 4),
                          "shade": round(k_ms.get(N.PROF_SHADE, 0.0), 4)},
           "pixels": n, "reflections": True}
    log("render: %.3f ms/frame (primary %.3f, shadow %.3f, reflect %.3f, shade %.3f), %.0f Mpixel/s" %
        (ms, res["kernels_ms"]["primary"], res["kernels_ms"]["shadow"], res["kernels_ms"]["reflect"],
         res["kernels_ms"]["shade"], res["mpix_s"]))
    if not args.no_cpu:
        import oracle
        img = d_rgba.cpu().numpy().reshape(-1, 4)
        rays = d_rays.cpu().numpy().view(N.RAY_DTYPE)
        pick = np.arange(n)
        threads = args.cpu_threads or min(os.cpu_count() or 1, 64)
        ow = oracle.World.from_scene(scene)
        dens = ow.density()
        t1 = time.time()
        exp = ow.render(ur, lights, rays[pick], threads=threads, density=dens)
        cpu_s = time.time() - t1
        res["bit_exact_sample"] = bool(np.array_equal(img[pick].view(np.uint32), exp.view(np.uint32)))
        res["cpu_oracle"] = {"mpix_s": round(pick.size / cpu_s / 1e6, 4), "threads": threads,
                             "sample": "the whole frame (%d pixels), %.2f s" % (pick.size, cpu_s)}
        log("render bit-exact vs oracle on %d pixels: %s; oracle %.3f Mpixel/s on %d threads" %
            (pick.size, res["bit_exact_sample"], res["cpu_oracle"]["mpix_s"], threads))
    gw.set_uniforms(u)
    return res


def scene_leg(args, scene):
    """The scene build of the 1280x352x960 grid from its 2002 SDF primitives
    on the GPU (px_create_from_prims: pairs, sort, per-brick replay, pool
    fill), wall time of the whole call including context creation, against
    the host scene builder (all host threads) and the oracle's dense
    rasteriser (one thread)."""
    from pixels_amd import GpuWorld
    prims = scene.prims()
    walls = []
    for i in range(4):
        t1 = time.time()
        gw = GpuWorld(None, prims=prims, dims=scene.dims, mats=scene.mats)
        gw.sync()
        walls.append(time.time() - t1)
        if i == 0:
            grid, sup = gw.read_grid()
            pool = gw.read_pool(0, scene.pool_used).view(np.uint16).reshape(-1)
            exact = bool(np.array_equal(grid, scene.grid) and np.array_equal(sup, scene.super_grid)
                         and np.array_equal(pool, scene.pool))
        gw.close()
    t1 = time.time()
    from pixels_amd.scene import LARGE, Scene
    Scene.build(LARGE, seed=args.seed)
    host_s = time.time() - t1
    res = {"gpu_ms": round(float(np.median(walls[1:])) * 1e3, 2), "gpu_ms_first": round(walls[0] * 1e3, 2),
           "host_builder_ms": round(host_s * 1e3, 2), "host_builder_threads": os.cpu_count(),
           "primitives": int(prims.shape[0]), "pool_bricks": scene.pool_used, "equals_host_builder": exact}
    log("scene build on the GPU: %.1f ms (first %.1f ms), host builder %.1f ms, equal: %s" %
        (res["gpu_ms"], res["gpu_ms_first"], res["host_builder_ms"], exact))
    if not args.no_cpu:
        import oracle
        t1 = time.time()
        vox, kind = oracle.raster(scene.dims, prims)
        res["cpu_oracle_ms"] = round((time.time() - t1) * 1e3, 1)
        del vox, kind
    return res


def extraction_leg(args, scene, gw, sim, impacts):
    """After the island pass: the extraction of every unsupported piece
    (chunks into the arena, crumbs and tumbling rubble into a 2^18 ring),
    then frames of falling chunks + debris + voxel writes, 60 fps."""
    from pixels_amd import CTRL
    from pixels_amd import _native as N
    vs = np.float32(1.0 / 80.0)
    u = np.zeros(N.PX_UNIFORM_FLOATS, np.float32)
    u[0:3] = scene.dims
    u[3], u[24], u[25], u[29] = vs, np.float32(1.0 / 60.0), np.float32(9.81) / vs, 5.0
    gw.set_uniforms(u)
    sim.particles_init(1 << 18)
    reps = []
    for _ in range(max(3, args.steps // 4)):
        gw.restore()
        sim.particles_init(1 << 18)
        sim.push_impacts(impacts)
        sim.encode(N.STAGE_FRAME_BEGIN | N.STAGE_EDIT)
        gw.sync()
        gw.profile_begin()
        sim.encode(N.STAGE_RESOLVE)
        prof = gw.profile_end()
        reps.append(prof[N.PROF_CHUNKS][0])
    ctrl, _, _ = sim.read_ctrl()
    res = {"resolve_ms": round(float(np.median(reps)), 4), "reps": len(reps),
           "chunks": int(ctrl[CTRL["S_ISLANDS"]]), "arena_voxels": int(ctrl[CTRL["C_ARENA"]]),
           "debris_spawned": int(ctrl[CTRL["C_SPAWNED"]])}
    frames = {"chunks": [], "particles": [], "writes": [], "frame": []}
    for f in range(6, 6 + max(20, args.steps)):
        u[29] = f
        gw.set_uniforms(u)
        gw.profile_begin()
        sim.encode(N.STAGE_FRAME_BEGIN | N.STAGE_CHUNKS | N.STAGE_PARTICLES | N.STAGE_WRITES)
        prof = gw.profile_end()
        c = prof.get(N.PROF_CHUNKS, (0.0, 0))[0]
        p = prof.get(N.PROF_PARTICLES, (0.0, 0))[0]
        w = prof.get(N.PROF_WRITES, (0.0, 0))[0]
        frames["chunks"].append(c)
        frames["particles"].append(p)
        frames["writes"].append(w)
        frames["frame"].append(c + p + w)
    ctrl, _, _ = sim.read_ctrl()
    res.update({k + "_ms": round(float(np.median(v)), 4) for k, v in frames.items()})
    res["frames"] = len(frames["frame"])
    res["falling_after"] = int(ctrl[CTRL["S_FALLING"]])
    log("extraction: %.3f ms (%d chunks, %d arena voxels, %d debris); frame of chunks + debris + writes %.3f ms"
        % (res["resolve_ms"], res["chunks"], res["arena_voxels"], res["debris_spawned"], res["frame_ms"]))
    return res


DEBRIS_RING = 1 << 18   # MAX_PARTICLES default, config.ts:38
DEBRIS_BIG = 1 << 21    # its maximum
# algorithmic bytes per live slot of k_particle_sim: the 32-byte record read
# and written back, the status byte; per tile of 1024 slots 8 bytes of counts
DEBRIS_SLOT_BYTES = 65


def surface_debris(dims, prim_rays, prim_hits, n, seed):
    """n debris records at the air voxels in front of primary-ray hits (where
    carveEffects spawns them, sim.wgsl:495-509), thrown off the surface at up
    to 60 voxels/s; flags and life as pixels_amd.scene.debris."""
    from pixels_amd import _native as N
    from pixels_amd.scene import debris
    rng = np.random.default_rng(seed)
    ok = np.nonzero(prim_hits["face"] < 6)[0]
    pick = ok[rng.integers(0, ok.size, n)]
    h = prim_hits[pick]
    dx, dy = dims[0], dims[1]
    key = h["key"].astype(np.int64)
    cell = np.stack([key % dx, (key // dx) % dy, key // (dx * dy)], 1)
    normal = np.zeros((n, 3), np.int64)
    ax = h["face"].astype(np.int64) // 2
    sgn = np.where(h["face"] % 2 == 0, 1, -1)  # dirOf order: +x, -x, +y, -y, +z, -z
    normal[np.arange(n), ax] = sgn
    rec = debris(dims, n, seed=seed, speed=20.0)
    rec["pos"] = (cell + normal + 0.5).astype(np.float32)
    rec["vel"] = (rec["vel"] + normal * rng.uniform(5.0, 40.0, (n, 1))).astype(np.float32)
    return rec.astype(N.PARTICLE_DTYPE)


def debris_leg(args, scene, gw, u, prim_rays, prim_hits):
    """Debris step (frameBegin + particleSim + alive / write lists) over a ring
    whose every slot holds a live particle thrown off the frame's visible
    surfaces, on the same grid as the frame; 60 fps, gravity 9.81 m/s^2 at 80
    voxels per metre."""
    import torch

    import oracle
    from pixels_amd import Simulation
    from pixels_amd import _native as N

    sim = Simulation(gw)
    uu = np.array(u, np.float32).copy()
    dt, vs = np.float32(1.0 / 60.0), np.float32(uu[3])
    g = np.float32(9.81) / vs
    uu[24], uu[25] = dt, g
    gw.set_uniforms(uu)
    res = {}
    for ring in (DEBRIS_RING, DEBRIS_BIG):
        lim = gw.limits
        recs = surface_debris(scene.dims, prim_rays, prim_hits, ring, seed=17)
        sim.particles_init(ring)
        d_recs = torch.from_numpy(recs.view(np.uint8).copy()).cuda()
        torch.cuda.synchronize()
        # fill the ring in budget-sized spawns, one frame each
        b = int(lim.spawn_budget)
        for off in range(0, ring, b):
            sim.frame_begin()
            sim.spawn_device(d_recs[off * 32:], min(b, ring - off))
        gw.sync()
        start = sim.read_particles()
        steps = max(args.steps, 10)
        for _ in range(args.warmup):
            sim.frame_begin()
            sim.particle_step()
        gw.sync()
        # timed: restart from the filled ring so every run sees the same states
        sim.particles_init(ring)
        for off in range(0, ring, b):
            sim.frame_begin()
            sim.spawn_device(d_recs[off * 32:], min(b, ring - off))
        gw.sync()
        gw.profile_begin()
        for _ in range(steps):
            sim.frame_begin()
            sim.particle_step()
        prof = gw.profile_end()
        alive = len(sim.read_alive())
        ms = prof[N.PROF_PARTICLES][0] / prof[N.PROF_PARTICLES][1]
        sim_ms = prof[N.PROF_PARTICLE_SIM][0] / prof[N.PROF_PARTICLE_SIM][1]
        algo = DEBRIS_SLOT_BYTES * ring + 8 * (ring // 1024)
        traffic, _ = profiled_traffic("k_particle_sim@%d" % ring, DEBRIS_TRAFFIC_FILE)
        r = {"ring": ring, "steps": steps, "ms_per_step": round(ms, 5), "sim_kernel_ms": round(sim_ms, 5),
             "particles_per_s": round(ring / (ms * 1e-3), 1), "alive_after": alive,
             "roofline": {"bound": "hbm", "kernel": "k_particle_sim", "algorithmic_bytes": algo,
                          "achieved": round(algo / (sim_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(algo / (sim_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": traffic,
                          "traffic_source": DEBRIS_TRAFFIC_FILE if traffic is not None else None}}
        log("debris ring %d: %.4f ms/step (sim kernel %.4f ms), %.3g particles/s, %d alive after %d steps" %
            (ring, ms, sim_ms, r["particles_per_s"], alive, steps))
        if ring == DEBRIS_RING and not args.no_cpu:
            # parity of one step from the filled ring, and the CPU baseline
            w = oracle.World.from_scene(scene)
            P = oracle.Particles(ring, spawn_budget=int(lim.spawn_budget), max_writes=int(lim.max_writes))
            sim.particles_init(ring)
            for off in range(0, ring, b):
                sim.frame_begin()
                P.frame_begin()
                sim.spawn_device(d_recs[off * 32:], min(b, ring - off))
                P.spawn(recs[off:off + b])
            sim.frame_begin()
            sim.particle_step()
            # CPU baseline: the same `steps`-step sequence from the filled
            # ring, repeated from a saved copy for about 10 s
            saved = (P.parts.copy(), P.ctr.copy())
            t1 = time.time()
            k = 0
            while k == 0 or (time.time() - t1 < 10.0 and k < 50 * steps):
                P.parts[:], P.ctr[:] = saved
                for _ in range(steps):
                    P.frame_begin()
                    P.step(w, dt, g, vs)
                    k += 1
            cpu_s = (time.time() - t1) / k
            got = sim.read_particles()
            P1 = oracle.Particles(ring, spawn_budget=int(lim.spawn_budget), max_writes=int(lim.max_writes))
            for off in range(0, ring, b):
                P1.frame_begin()
                P1.spawn(recs[off:off + b])
            P1.frame_begin()
            P1.step(w, dt, g, vs)
            r["bit_exact_vs_oracle"] = bool(got.tobytes() == P1.parts.tobytes() and
                                            np.array_equal(sim.read_alive(), P1.alive_list()) and
                                            np.array_equal(sim.read_writes(), P1.write_list()))
            r["cpu_baseline"] = {"value": round(ring / cpu_s, 1), "unit": "particles/s", "cores": 1, "kind": "port",
                                 "sample": "%d oracle steps of the same %d-slot ring, %.3f s each" % (k, ring, cpu_s)}
            log("debris bit-exact vs oracle: %s; oracle %.3f s/step" % (r["bit_exact_vs_oracle"], cpu_s))
        res["ring_%d" % ring] = r
        del d_recs
    return res


def batch_leg(args):
    """Queries of very uneven counts on the 2560x704x1920 scene (the
    full-scale generator scaled x2, 160 voxels per metre): 2M primary, 3M
    shadow, 50k shoot rays from 64 origins, 262,144 debris segments of
    length U[1,200] voxels.  One autotuned persistent launch replayed as a HIP
    graph, against the four per-kind launches; outputs compared bit for bit,
    and a sample of each kind against the oracle."""
    import torch
    from pixels_amd import GpuWorld, Scene, Simulation
    from pixels_amd import _native as N
    from pixels_amd.scene import LARGE_SCALED
    from pixels_amd.workloads import mixed_queries

    t0 = time.time()
    scene = Scene.build(LARGE_SCALED, seed=args.seed, dims=(2560, 704, 1920), scale=2.0)
    vs = 1.0 / 160.0
    log("batch scene %s fill %.3f pool bricks %d (%.1f%% of cells) built in %.1fs" %
        (scene.dims, scene.fill, scene.pool_used, 100.0 * scene.pool_used / scene.grid.size, time.time() - t0))
    gw = GpuWorld(scene, voxel_size=vs)
    sim = Simulation(gw)
    dev = torch.device("cuda:0")
    host = mixed_queries(scene, sim, gw, args.seed, vs=vs)
    rays = {k: torch.from_numpy(np.ascontiguousarray(v).view(np.uint8).copy()).to(dev) for k, v in host.items()}
    outb = {k: torch.empty(v.shape[0] * 16, dtype=torch.uint8, device=dev) for k, v in host.items()}
    outp = {k: torch.empty(v.shape[0] * 16, dtype=torch.uint8, device=dev) for k, v in host.items()}
    counts = {k: int(v.shape[0]) for k, v in host.items()}
    total = sum(counts.values())
    q = sim.query_set({k: (rays[k], counts[k]) for k in host}, outb)
    torch.cuda.synchronize()

    def per_kind():
        for k in host:
            sim.trace_device(k, rays[k], counts[k], outp[k])

    reps = max(3, args.steps // 2)
    per_kind()
    gw.sync()
    gw.profile_begin()
    for _ in range(reps):
        per_kind()
    plain_ms = gw.profile_end()[N.PROF_TRACE][0] / reps
    shape, tuned_ms = sim.autotune_batch(q, repeats=2)
    sim.trace_batch(q)
    gw.sync()
    gw.profile_begin()
    for _ in range(reps):
        sim.trace_batch(q)
    batch_ms = gw.profile_end()[N.PROF_BATCH][0] / reps
    same = all(np.array_equal(outb[k].cpu().numpy(), outp[k].cpu().numpy()) for k in host)
    res = {"counts": {"primary": counts[0], "shadow": counts[1], "shoot": counts[2], "segment": counts[3]},
           "grid": list(scene.dims), "pool_bricks": scene.pool_used,
           "nonempty_brick_frac": round(float((scene.grid != 0).mean()), 4),
           "batch_ms": round(batch_ms, 4), "per_kind_ms": round(plain_ms, 4),
           "batch_grays_s": round(total / (batch_ms * 1e-3) / 1e9, 4),
           "per_kind_grays_s": round(total / (plain_ms * 1e-3) / 1e9, 4),
           "speedup": round(plain_ms / batch_ms, 3), "launch_shape": shape,
           "autotune_best_ms": round(tuned_ms, 4), "bit_exact_vs_per_kind": bool(same)}
    log("batched %d queries: graph %.3f ms vs per-kind launches %.3f ms (x%.2f), shape %s, same bits %s" %
        (total, batch_ms, plain_ms, plain_ms / batch_ms, shape, same))
    if not args.no_cpu:
        import oracle
        ow = oracle.World.from_scene(scene)
        ok = True
        for k in host:
            idx = np.linspace(0, counts[k] - 1, min(4000, counts[k])).astype(np.int64)
            got = outb[k].cpu().numpy().view(N.HIT_DTYPE)[idx]
            exp = ow.trace(k, host[k][idx])
            ok = ok and bool(np.array_equal(got.view(np.uint8), exp.view(np.uint8)))
        res["bit_exact_vs_oracle_sample"] = ok
        log("batched sample bit-exact vs oracle: %s" % ok)
    if not args.no_learned:
        res["learned"] = batch_learned(args, scene, sim, gw, host, rays, counts, q, outb, plain_ms, batch_ms)
    gw.close()
    return res


def batch_learned(args, scene, sim, gw, host, rays, counts, q, outb, per_kind_ms, batch_ms):
    """The learned-binned batched launch on the 2560x704x1920 mix.  The 3M
    shadow queries arrive in the order fsMain makes them (pixel, then sun and
    lights), which is the plain launches' order.  A 16-64-64-1 network is
    trained on this scene (Adam, lr 1e-3, batch 65,536, 12 epochs) on the
    shadow rays' step counts from 128 seeded 480x270 cameras, then orders the
    measured shadow queries by (direction face, one of 32 predicted-cost
    buckets) before one px_trace_batch_perm launch; primary rays (8x8 tiles),
    shoot rays and segments stay in index order.  Against: the per-kind
    launches and the batched launch in index order, and the same sort with
    one cost bucket (direction face only, no network)."""
    import torch
    from pixels_amd import _native as N
    from pixels_amd.learned import CostModel, log2_steps, wave_cost
    from pixels_amd.workloads import shadow_queries

    dev = torch.device("cuda:0")
    vs = 1.0 / 160.0
    TW, TH, CAMS = 480, 270, 128
    t0 = time.time()
    model = CostModel(gw, seed=args.seed)
    feats, targets = [], []
    cam = 5000
    while len(targets) < CAMS and cam < 5000 + 4 * CAMS:
        cam += 1
        prim = scene.camera_rays(TW, TH, 90.0, seed=cam, tile=8)
        dp = torch.from_numpy(prim.view(np.uint8).copy()).to(dev)
        op = torch.empty(prim.shape[0] * 16, dtype=torch.uint8, device=dev)
        sim.trace_device(0, dp, prim.shape[0], op)
        gw.sync()
        u, lights = scene.frame_inputs(TW, TH, seed=cam, voxel_size=vs, light_radius_vox=8192.0)
        try:
            sh = shadow_queries(scene, prim, op.cpu().numpy().view(N.HIT_DTYPE), u, lights, TW * TH)
        except IndexError:
            continue
        ds = torch.from_numpy(sh.view(np.uint8).copy()).to(dev)
        n = sh.shape[0]
        f = torch.empty(n * 16, dtype=torch.int16, device=dev)
        model.features(1, ds, n, f)
        o = torch.empty(n * 16, dtype=torch.uint8, device=dev)
        sim.trace_device(1, ds, n, o)
        gw.sync()
        feats.append(f)
        targets.append(torch.from_numpy(log2_steps(o.cpu().numpy().view(N.HIT_DTYPE)["steps"])).to(dev))
    tf, tt = torch.cat(feats), torch.cat(targets)
    del feats, targets
    n_train = tt.numel()
    order = torch.randperm(n_train, device=dev, generator=torch.Generator(device=dev).manual_seed(args.seed))
    tf = tf.view(n_train, 16)[order].reshape(-1).contiguous()
    tt = tt[order].contiguous()
    bs, epochs = 65536, 12
    first = model.train_step(tf, tt, bs, lr=1e-3, want_loss=True)
    for _ in range(epochs):
        for off in range(0, n_train - bs + 1, bs):
            model.train_step(tf[off * 16:], tt[off:], bs, lr=1e-3)
    last = model.train_step(tf, tt, bs, lr=1e-3, want_loss=True)
    del tf, tt
    train_s = time.time() - t0

    n_sh = counts[1]
    perm = torch.empty(n_sh, dtype=torch.int32, device=dev)
    outl = {k: torch.empty_like(v) for k, v in outb.items()}
    ql = sim.query_set({k: (rays[k], counts[k]) for k in host}, outl)

    def learned(buckets=32, wlog=WINDOW_LOG2):
        model.sort_rays(1, rays[1], n_sh, perm, buckets=buckets, lo=0.0, hi=9.0, faces=True, window_log2=wlog)
        sim.trace_batch_perm(ql, {1: perm})

    # each order with the launch shape the autotuner picks for it
    shapes = {"plain": sim.get_batch_shape()}
    for name, buckets, wlog in (("faces", 1, WINDOW_LOG2), ("global", 32, 0), ("learned", 32, WINDOW_LOG2)):
        learned(buckets, wlog)
        shapes[name], _ = sim.autotune_batch(ql, repeats=2, perms={1: perm})
    use = lambda name: sim.set_batch_shape(shapes[name]["block"], shapes[name]["blocks"], shapes[name]["chunk"])
    learned()
    gw.sync()
    same = all(torch.equal(outl[k], outb[k]) for k in host)
    steps = outb[1].cpu().numpy().view(N.HIT_DTYPE)["steps"].astype(np.int64)
    pm = perm.cpu().numpy().astype(np.int64)
    wc_plain, wc_learned = wave_cost(steps), wave_cost(steps[pm])
    reps, rounds = 3, 7
    pb, lb, ls, fb, fs, gb, gs = [], [], [], [], [], [], []
    for _ in range(rounds):  # interleaved, medians
        use("plain")
        sim.trace_batch(q)
        gw.profile_begin()
        for _ in range(reps):
            sim.trace_batch(q)
        pb.append(gw.profile_end()[N.PROF_BATCH][0] / reps)
        for name, buckets, wlog, tb, ts in (("learned", 32, WINDOW_LOG2, lb, ls), ("faces", 1, WINDOW_LOG2, fb, fs),
                                            ("global", 32, 0, gb, gs)):
            use(name)
            learned(buckets, wlog)
            gw.profile_begin()
            for _ in range(reps):
                learned(buckets, wlog)
            p = gw.profile_end()
            tb.append(p[N.PROF_BATCH][0] / reps)
            ts.append(p.get(N.PROF_MLP, (0.0, 0))[0] / reps)
    learned(1)
    gw.sync()
    wc_faces = wave_cost(steps[perm.cpu().numpy().astype(np.int64)])
    learned(32, 0)
    gw.sync()
    wc_global = wave_cost(steps[perm.cpu().numpy().astype(np.int64)])
    use("learned")
    learned()
    gw.sync()
    same = same and all(torch.equal(outl[k], outb[k]) for k in host)
    use("plain")
    med = lambda x: float(np.median(x))
    p_ms, l_ms, l_sort, f_ms, f_sort = med(pb), med(lb), med(ls), med(fb), med(fs)
    g_ms = med(gb) + med(gs)
    total = sum(counts.values())
    r = {"shadow_order": "pixel, then sun and lights (fsMain's order)",
         "learned_order": "(direction face, 32 predicted-cost buckets) inside each window of 2^%d queries" % WINDOW_LOG2,
         "train_cameras": CAMS,
         "train_pairs": int(n_train), "epochs": epochs, "batch": bs, "loss_first": round(first, 4),
         "loss_last": round(last, 4), "train_wall_s": round(train_s, 2),
         "per_kind_ms": round(per_kind_ms, 4), "batch_ms": round(p_ms, 4),
         "learned_trace_ms": round(l_ms, 4), "learned_sort_ms": round(l_sort, 4),
         "learned_ms": round(l_ms + l_sort, 4), "learned_grays_s": round(total / ((l_ms + l_sort) * 1e-3) / 1e9, 4),
         "speedup_vs_per_kind": round(per_kind_ms / (l_ms + l_sort), 3),
         "speedup_vs_batch": round(p_ms / (l_ms + l_sort), 3),
         "faces_only_ms": round(f_ms + f_sort, 4), "faces_only_speedup_vs_per_kind": round(per_kind_ms / (f_ms + f_sort), 3),
         "global_learned_ms": round(g_ms, 4), "global_learned_speedup_vs_per_kind": round(per_kind_ms / g_ms, 3),
         "launch_shapes": shapes, "shadow_wave_cost": {"index_order": wc_plain, "learned": wc_learned, "faces_only": wc_faces,
                              "global_learned": wc_global},
         "bit_identical": bool(same)}
    log("learned-binned batch: %.3f + %.3f ms sort vs per-kind %.3f ms (x%.2f), vs batch %.3f ms (x%.2f); "
        "faces only %.3f ms (x%.2f); one global sort %.3f ms; same bits %s" %
        (l_ms, l_sort, per_kind_ms, r["speedup_vs_per_kind"], p_ms, r["speedup_vs_batch"], f_ms + f_sort,
         r["faces_only_speedup_vs_per_kind"], g_ms, same))
    model.close()
    return r


def learned_leg(args):
    """Learned ray-cost binning on the 512x256x512 scenes (plain, and with
    the structural floor) and on the 1280x352x960 frame scene.  Per scene
    and query kind (primary, shadow), a 16-64-64-1 BF16 MLP is trained on
    the GPU with Adam (lr 1e-3, batch 65,536, 12 epochs) on the 16.6M
    (ray, step count) pairs that the plain DDA produces for 128 seeded
    480x270 cameras, then used on a held-out 1920x1080 camera: features +
    prediction + a stable global sort by (direction face, one of 32 cost
    buckets) + the permuted launch, against the unsorted launch of the same
    rays; with the cost key alone and the face alone as ablations.  Shadow
    rays are compared in two plain orders: pixel then slot, and the frame's
    own queue order (64-pixel region, slot, pixel).  On the frame scene the
    held-out camera's every cast shadow ray (12.1M) is also timed through the
    frame launch itself, with the predictor's order inside each region, and
    the primary + shadow rays through the batched launch with and without
    the learned orders."""
    import torch
    from pixels_amd import GpuWorld, Scene, Simulation
    from pixels_amd import _native as N
    from pixels_amd.learned import CostModel, bin_rays, log2_steps, trace_perm, wave_cost
    from pixels_amd.scene import LARGE, SMALL, SMALL_FLOOR
    from pixels_amd.workloads import frame_order, region_wave_cost, shadow_queries

    dev = torch.device("cuda:0")
    W, H = 1920, 1080
    TW, TH, N_TRAIN_CAMS = 480, 270, 128
    out = {"scenes": [], "train_pairs": 0}

    def tr(sim, gw, kind, rays_np, d_rays):
        n = rays_np.shape[0]
        o = torch.empty(n * 16, dtype=torch.uint8, device=dev)
        sim.trace_device(kind, d_rays, n, o)
        gw.sync()
        return o

    def prof_ms(gw, fn, slots, reps):
        gw.profile_begin()
        for _ in range(reps):
            fn()
        p = gw.profile_end()
        return sum(p.get(k, (0.0, 0))[0] for k in slots) / reps

    t_all = time.time()
    for kind_scene, seed in ((SMALL, 1), (SMALL_FLOOR, 3), (LARGE, 2)):
        scene = Scene.build(kind_scene, seed=seed)
        gw = GpuWorld(scene)
        sim = Simulation(gw)
        frame_scene = kind_scene is LARGE

        def rays_of(cam, w, h, all_shadow=False):
            prim = scene.camera_rays(w, h, 90.0, seed=cam, tile=8)
            dp = torch.from_numpy(prim.view(np.uint8).copy()).to(dev)
            ph = tr(sim, gw, 0, prim, dp).cpu().numpy().view(N.HIT_DTYPE)
            u, lights = scene.frame_inputs(w, h, seed=cam)
            sh, pix, slot = shadow_queries(scene, prim, ph, u, lights, None if all_shadow else w * h, codes=True)
            ds = torch.from_numpy(sh.view(np.uint8).copy()).to(dev)
            return {0: (prim, dp), 1: (sh, ds)}, (pix, slot, u, lights)

        # training cameras: 128 seeded 480x270 views (many viewpoints, so the
        # network learns the scene rather than a few cameras; seeds from
        # 1000 x scene seed + 1); one 1920x1080 validation camera picks
        # between initialisations; the held-out 1920x1080 camera is measured
        # on the frame scene the held-out camera is the measured frame's own
        # (seed 2: 2,073,600 primary and 12.1M shadow rays)
        held_out = args.seed if frame_scene else seed * 1000 + 999
        data = {0: [], 1: []}
        probe = CostModel(gw, seed=args.seed)
        cams, cam = 0, seed * 1000
        while cams < N_TRAIN_CAMS and cam < seed * 1000 + 4 * N_TRAIN_CAMS:
            cam += 1
            try:
                rk, _ = rays_of(cam, TW, TH)
            except IndexError:  # no shadow rays from this camera
                continue
            cams += 1
            for kind, (r, d) in rk.items():
                n = r.shape[0]
                f = torch.empty(n * 16, dtype=torch.int16, device=dev)
                probe.features(kind, d, n, f)
                steps = tr(sim, gw, kind, r, d).cpu().numpy().view(N.HIT_DTYPE)["steps"]
                data[kind].append((f, torch.from_numpy(log2_steps(steps)).to(dev)))
        probe.close()
        val, _ = rays_of(seed * 1000 + 998, W, H)
        res = {"scene": list(scene.dims), "train_cameras": cams, "train_camera_size": [TW, TH]}
        held, (h_pix, h_slot, h_u, h_lights) = rays_of(held_out, W, H, all_shadow=frame_scene)
        models = {}
        for kind in (0, 1):
            # one network per query kind, trained on that kind's pairs; of
            # four initialisations, the one whose binned launch is fastest on
            # the validation camera is used on the held-out camera
            tf = torch.cat([f for f, _ in data[kind]])
            tt = torch.cat([t for _, t in data[kind]])
            n_train = tt.numel()
            out["train_pairs"] += n_train
            order = torch.randperm(n_train, device=dev,
                                   generator=torch.Generator(device=dev).manual_seed(args.seed))
            tf = tf.view(n_train, 16)[order].reshape(-1).contiguous()
            tt = tt[order].contiguous()
            bs, epochs = 65536, 12
            vr, vd = val[kind]
            nv = vr.shape[0]
            v_perm = torch.empty(nv, dtype=torch.int32, device=dev)
            v_out = torch.empty(nv * 16, dtype=torch.uint8, device=dev)
            cands = []
            for init in range(4):
                model = CostModel(gw, seed=args.seed + 1000 * init)
                gw.sync()
                ext = torch.cuda.ExternalStream(gw.stream)
                ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                t0 = time.time()
                first = model.train_step(tf, tt, bs, lr=1e-3, want_loss=True)
                ev0.record(ext)
                for ep in range(epochs):
                    for off in range(0, n_train - bs + 1, bs):
                        model.train_step(tf[off * 16:], tt[off:], bs, lr=1e-3)
                ev1.record(ext)
                last = model.train_step(tf, tt, bs, lr=1e-3, want_loss=True)
                train_s = time.time() - t0
                train_gpu_ms = ev0.elapsed_time(ev1)
                model.trace_sorted(kind, vd, nv, v_perm, v_out, faces=True)
                gw.sync()
                v_ms = prof_ms(gw, lambda: model.trace_sorted(kind, vd, nv, v_perm, v_out, faces=True),
                               (N.PROF_TRACE, N.PROF_MLP), 3)
                cands.append((v_ms, init, model, first, last, train_s, train_gpu_ms))
            cands.sort(key=lambda c: c[0])
            v_ms, init, model, first, last, train_s, train_gpu_ms = cands[0]
            for c in cands[1:]:
                c[2].close()
            del tf, tt
            r, d = held[kind]
            n = r.shape[0]
            plain = torch.empty(n * 16, dtype=torch.uint8, device=dev)
            binned = torch.empty(n * 16, dtype=torch.uint8, device=dev)
            f = torch.empty(n * 16, dtype=torch.int16, device=dev)
            pred = torch.empty(n, dtype=torch.float32, device=dev)
            perm = torch.empty(n, dtype=torch.int32, device=dev)

            def run_plain(dd=d):
                sim.trace_device(kind, dd, n, plain)

            def run_binned(faces=True, buckets=32):
                model.trace_sorted(kind, d, n, perm, binned, buckets=buckets, lo=0.0, hi=9.0, faces=faces)
            run_plain()
            run_binned()
            gw.sync()
            model.features(kind, d, n, f)
            model.predict(f, n, pred)
            ref_perm = torch.empty(n, dtype=torch.int32, device=dev)
            bin_rays(gw, pred, n, ref_perm, window=n, buckets=32, lo=0.0, hi=9.0)
            run_binned(faces=False)
            gw.sync()
            same_order = bool(torch.equal(ref_perm, perm))
            run_binned()
            gw.sync()
            hits = plain.cpu().numpy().view(N.HIT_DTYPE)
            steps = hits["steps"].astype(np.int64)
            pr = pred.cpu().numpy()
            pm = perm.cpu().numpy().astype(np.int64)
            # the frame's queue order of the same shadow rays, and the
            # predictor's order inside each of its regions (stable by bucket)
            orders = {}
            if kind == 1:
                fo = frame_order(h_pix, h_slot)
                bucket = np.clip(np.floor(pr * (32.0 / 9.0)), 0, 31).astype(np.int64)
                reg = h_pix // 64
                fr = fo[np.lexsort((np.arange(fo.size), bucket[fo], reg[fo]))]
                orders = {"frame": fo, "frame_region_binned": fr}
            d_orders = {k: torch.from_numpy(v.astype(np.int32)).to(dev) for k, v in orders.items()}
            # interleaved rounds, medians: an occasional slow round (seen
            # either way) does not decide the comparison
            reps, rounds = 3, 7
            p_r, t_r, s_r = [], [], []
            o_r = {k: [] for k in orders}
            abl = {"cost_only": ([], []), "faces_only": ([], [])}
            for _ in range(rounds):
                p_r.append(prof_ms(gw, run_plain, (N.PROF_TRACE,), reps))
                gw.profile_begin()
                for _ in range(reps):
                    run_binned()
                prof = gw.profile_end()
                t_r.append(prof[N.PROF_TRACE][0] / reps)
                s_r.append(prof.get(N.PROF_MLP, (0.0, 0))[0] / reps)
                for k, dperm in d_orders.items():
                    o_r[k].append(prof_ms(gw, lambda: trace_perm(gw, kind, d, dperm, n, plain), (N.PROF_TRACE,), reps))
                # ablations: the learned cost key alone, the direction alone
                for name, kw in (("cost_only", dict(faces=False)), ("faces_only", dict(buckets=1))):
                    gw.profile_begin()
                    for _ in range(reps):
                        run_binned(**kw)
                    prof = gw.profile_end()
                    abl[name][0].append(prof[N.PROF_TRACE][0] / reps)
                    abl[name][1].append(prof.get(N.PROF_MLP, (0.0, 0))[0] / reps)
            run_binned()  # leave the main order in perm / binned
            p_ms, b_trace_ms, sort_ms = float(np.median(p_r)), float(np.median(t_r)), float(np.median(s_r))
            b_ms = b_trace_ms + sort_ms
            ablation = {name: {"speedup": round(p_ms / (float(np.median(t)) + float(np.median(sr))), 3),
                               "speedup_trace_only": round(p_ms / float(np.median(t)), 3)}
                        for name, (t, sr) in abl.items()}
            rec = {
                "rays": n, "train_pairs": n_train, "epochs": epochs, "batch": bs, "inits_tried": len(cands),
                "validation_binned_ms": [round(c[0], 4) for c in sorted(cands, key=lambda c: c[1])],
                "init_chosen": init,
                "loss_first": round(first, 4), "loss_last": round(last, 4), "train_wall_s": round(train_s, 3),
                "train_gpu_ms": round(train_gpu_ms, 3),
                "held_out_log2_mse": round(float(np.mean((pr - log2_steps(steps)) ** 2)), 4),
                "plain_order": "pixel, then slot" if kind == 1 else "camera tiles of 8x8",
                "plain_grays_s": round(n / (p_ms * 1e-3) / 1e9, 4),
                "binned_grays_s": round(n / (b_ms * 1e-3) / 1e9, 4),
                "binned_trace_only_grays_s": round(n / (b_trace_ms * 1e-3) / 1e9, 4),
                "sort_ms": round(sort_ms, 4), "plain_ms": round(p_ms, 4), "binned_trace_ms": round(b_trace_ms, 4),
                "rounds": rounds, "plain_ms_range": [round(min(p_r), 4), round(max(p_r), 4)],
                "binned_trace_ms_range": [round(min(t_r), 4), round(max(t_r), 4)],
                "speedup": round(p_ms / b_ms, 3), "speedup_trace_only": round(p_ms / b_trace_ms, 3),
                "mae_steps": round(float(np.mean(np.abs(np.exp2(pr) - np.maximum(steps, 1)))), 2),
                "mae_steps_constant_median": round(float(np.mean(np.abs(np.median(np.maximum(steps, 1)) -
                                                                        np.maximum(steps, 1)))), 2),
                "mean_steps": round(float(steps.mean()), 2),
                "wave_cost_plain": wave_cost(steps), "wave_cost_binned": wave_cost(steps[pm]),
                "bit_identical": bool(torch.equal(plain, binned)), "order_equals_separate_kernels": same_order,
                "keys": "direction face x 32 cost buckets", "ablation": ablation}
            for k, o in orders.items():
                ms = float(np.median(o_r[k]))
                rec["order_" + k] = {"trace_ms": round(ms, 4), "grays_s": round(n / (ms * 1e-3) / 1e9, 4),
                                     "wave_cost": wave_cost(steps[o]),
                                     "region_wave_cost": region_wave_cost(steps[o], (h_pix // 64)[o])}
            if orders:
                rec["order_frame_region_binned"]["sort_ms_not_included"] = True
                # the ideal in-region order (true step counts) bounds what any predictor can give there
                reg = h_pix // 64
                fo = orders["frame"]
                ideal = fo[np.lexsort((np.arange(fo.size), -steps[fo], reg[fo]))]
                rec["region_wave_cost_true_cost_order"] = region_wave_cost(steps[ideal], reg[ideal])
            res["primary" if kind == 0 else "shadow"] = rec
            models[kind] = model
        if frame_scene:
            res["frame_and_batch"] = frame_and_batch_learned(args, scene, gw, sim, held, (h_pix, h_slot, h_u, h_lights),
                                                            models)
        for m in models.values():
            m.close()
        del data
        out["scenes"].append(res)
        log("learned binning %s: %s" % (scene.dims, json.dumps(res)))
        gw.close()
    out["wall_s"] = round(time.time() - t_all, 2)
    return out


def frame_and_batch_learned(args, scene, gw, sim, held, codes, models):
    """On the frame scene's held-out camera: its every cast shadow ray
    through the frame launch itself (k_frame_primary + k_trace_shadow_queue),
    and the primary + shadow rays through the batched launch in index order
    and in the learned orders (px_mlp_sort_rays per kind, then
    px_trace_batch_perm)."""
    import torch
    from pixels_amd import _native as N
    from pixels_amd.workloads import frame_order
    dev = torch.device("cuda:0")
    pix, slot, u, lights = codes
    (prim, d_prim), (sh, d_sh) = held[0], held[1]
    n = prim.shape[0]
    gw.set_uniforms(u)
    gw.set_lights(lights)
    sim.set_frame_layout(True)  # the measured frame's slot-planar output
    slots = sim.frame_slots()
    dp = torch.empty(n * 16, dtype=torch.uint8, device=dev)
    ds = torch.empty(n * slots * 16, dtype=torch.uint8, device=dev)
    nc = sim.trace_frame_device(d_prim, n, dp, ds, count=True)
    gw.sync()
    fr = []
    for _ in range(7):
        gw.profile_begin()
        for _ in range(3):
            sim.trace_frame_device(d_prim, n, dp, ds)
        p = gw.profile_end()
        fr.append((p[N.PROF_PRIMARY][0] / 3, p[N.PROF_SHADOW][0] / 3))
    f_prim, f_sh = float(np.median([a for a, _ in fr])), float(np.median([b for _, b in fr]))
    res = {"frame": {"shadow_rays": int(nc), "materialized_shadow_rays": int(sh.shape[0]),
                     "shadow_kernel_ms": round(f_sh, 4), "shadow_grays_s": round(nc / (f_sh * 1e-3) / 1e9, 4),
                     "primary_kernel_ms": round(f_prim, 4)}}
    # batched: the held-out primaries + the shadow rays in the frame's order
    fo = torch.from_numpy(frame_order(pix, slot).astype(np.int64)).to(dev)
    sh_rays = d_sh.view(-1, 32)[fo].reshape(-1).contiguous()
    ns = sh.shape[0]
    rays = {0: (d_prim, n), 1: (sh_rays, ns)}
    outs = {0: torch.empty(n * 16, dtype=torch.uint8, device=dev), 1: torch.empty(ns * 16, dtype=torch.uint8, device=dev)}
    outs_b = {0: torch.empty(n * 16, dtype=torch.uint8, device=dev), 1: torch.empty(ns * 16, dtype=torch.uint8, device=dev)}
    q = sim.query_set(rays, outs)
    qb = sim.query_set(rays, outs_b)
    perms = {k: torch.empty(rays[k][1], dtype=torch.int32, device=dev) for k in (0, 1)}

    def learned_batch():
        for k in (0, 1):
            models[k].sort_rays(k, rays[k][0], rays[k][1], perms[k], buckets=32, lo=0.0, hi=9.0, faces=True)
        sim.trace_batch_perm(qb, perms)
    sim.trace_batch(q)
    learned_batch()
    gw.sync()
    same = all(torch.equal(outs[k], outs_b[k]) for k in (0, 1))
    pb, lb, ls = [], [], []
    for _ in range(7):
        gw.profile_begin()
        for _ in range(3):
            sim.trace_batch(q)
        pb.append(gw.profile_end()[N.PROF_BATCH][0] / 3)
        gw.profile_begin()
        for _ in range(3):
            learned_batch()
        p = gw.profile_end()
        lb.append(p[N.PROF_BATCH][0] / 3)
        ls.append(p.get(N.PROF_MLP, (0.0, 0))[0] / 3)
    p_ms, l_ms, s_ms = float(np.median(pb)), float(np.median(lb)), float(np.median(ls))
    tot = n + ns
    res["batch"] = {"queries": {"primary": n, "shadow": int(ns)}, "shadow_order": "the frame's queue order",
                    "plain_ms": round(p_ms, 4), "learned_trace_ms": round(l_ms, 4), "learned_sort_ms": round(s_ms, 4),
                    "plain_grays_s": round(tot / (p_ms * 1e-3) / 1e9, 4),
                    "learned_grays_s": round(tot / ((l_ms + s_ms) * 1e-3) / 1e9, 4),
                    "speedup": round(p_ms / (l_ms + s_ms), 3), "speedup_trace_only": round(p_ms / l_ms, 3),
                    "bit_identical": bool(same)}
    del sh_rays, outs, outs_b, perms
    # the same queries with the shadow rays in the order fsMain makes them
    # (pixel, then sun and lights: the plain launches' order), the primary
    # rays in 8x8 tiles: per-kind launches, the batched launch, and the
    # batched launch with the shadow rays in the learned (direction face,
    # cost) order or in direction-face order alone (no network)
    rays = {0: (d_prim, n), 1: (d_sh, ns)}
    outs = {0: torch.empty(n * 16, dtype=torch.uint8, device=dev), 1: torch.empty(ns * 16, dtype=torch.uint8, device=dev)}
    outs_b = {0: torch.empty(n * 16, dtype=torch.uint8, device=dev), 1: torch.empty(ns * 16, dtype=torch.uint8, device=dev)}
    q = sim.query_set(rays, outs)
    qb = sim.query_set(rays, outs_b)
    perm = torch.empty(ns, dtype=torch.int32, device=dev)

    def per_kind():
        for k in (0, 1):
            sim.trace_device(k, rays[k][0], rays[k][1], outs[k])

    def binned(buckets, wlog=WINDOW_LOG2):
        models[1].sort_rays(1, d_sh, ns, perm, buckets=buckets, lo=0.0, hi=9.0, faces=True, window_log2=wlog)
        sim.trace_batch_perm(qb, {1: perm})
    # each batched launch below with the shape the autotuner picks for its order
    shapes = {}
    shapes["plain"], _ = sim.autotune_batch(q, repeats=2)
    for name, buckets, wlog in (("faces", 1, WINDOW_LOG2), ("global", 32, 0), ("learned", 32, WINDOW_LOG2)):
        binned(buckets, wlog)
        shapes[name], _ = sim.autotune_batch(qb, repeats=2, perms={1: perm})
    use = lambda name: sim.set_batch_shape(shapes[name]["block"], shapes[name]["blocks"], shapes[name]["chunk"])
    use("plain")
    sim.trace_batch(q)
    use("learned")
    binned(32)
    gw.sync()
    same = all(torch.equal(outs[k], outs_b[k]) for k in (0, 1))
    kb, pb, lb, ls, fb, fs, gb, gs = [], [], [], [], [], [], [], []
    for _ in range(7):
        gw.profile_begin()
        for _ in range(3):
            per_kind()
        kb.append(gw.profile_end()[N.PROF_TRACE][0] / 3)
        use("plain")
        sim.trace_batch(q)
        gw.profile_begin()
        for _ in range(3):
            sim.trace_batch(q)
        pb.append(gw.profile_end()[N.PROF_BATCH][0] / 3)
        for name, buckets, wlog, tb, ts in (("learned", 32, WINDOW_LOG2, lb, ls), ("faces", 1, WINDOW_LOG2, fb, fs),
                                            ("global", 32, 0, gb, gs)):
            use(name)
            binned(buckets, wlog)
            gw.profile_begin()
            for _ in range(3):
                binned(buckets, wlog)
            p = gw.profile_end()
            tb.append(p[N.PROF_BATCH][0] / 3)
            ts.append(p.get(N.PROF_MLP, (0.0, 0))[0] / 3)
    same = same and all(torch.equal(outs[k], outs_b[k]) for k in (0, 1))
    med = lambda x: float(np.median(x))
    k_ms, p_ms, l_ms, s_ms, f_ms = med(kb), med(pb), med(lb), med(ls), med(fb) + med(fs)
    g_ms = med(gb) + med(gs)
    res["batch_fsmain_order"] = {
        "queries": {"primary": n, "shadow": int(ns)}, "shadow_order": "pixel, then sun and lights (fsMain's order)",
        "learned_order": "(direction face, 32 predicted-cost buckets) inside each window of 2^%d queries" % WINDOW_LOG2,
        "per_kind_ms": round(k_ms, 4), "batch_ms": round(p_ms, 4), "learned_trace_ms": round(l_ms, 4),
        "learned_sort_ms": round(s_ms, 4), "learned_grays_s": round(tot / ((l_ms + s_ms) * 1e-3) / 1e9, 4),
        "per_kind_grays_s": round(tot / (k_ms * 1e-3) / 1e9, 4),
        "speedup_vs_per_kind": round(k_ms / (l_ms + s_ms), 3), "speedup_vs_batch": round(p_ms / (l_ms + s_ms), 3),
        "faces_only_ms": round(f_ms, 4), "faces_only_speedup_vs_per_kind": round(k_ms / f_ms, 3),
        "global_learned_ms": round(g_ms, 4), "global_learned_speedup_vs_per_kind": round(k_ms / g_ms, 3),
        "launch_shapes": shapes, "bit_identical": bool(same)}
    log("frame scene, fsMain order: learned-binned batch %.3f + %.3f ms vs per-kind %.3f ms (x%.2f), "
        "vs batch %.3f ms (x%.2f); faces only %.3f ms; one global sort %.3f ms; same bits %s" %
        (l_ms, s_ms, k_ms, k_ms / (l_ms + s_ms), p_ms, p_ms / (l_ms + s_ms), f_ms, g_ms, same))
    return res

if __name__ == "__main__":
    sys.exit(main())
