"""Synthetic query sets of the measured workloads, built on the host (they
are inputs: rays, not results).  bench.py and the GPU tests build them here
so that both trace exactly the same rays.

  shadow_queries  the shadow rays of shade() (raymarch.wgsl:323-356) for a
                  set of primary hits, one ray per (pixel, slot) in
                  pixel-then-slot order
  mixed_queries   the uneven batch of the 2560x704x1920 scene: 2M primary
                  rays (a 2048x1024 camera), 3M shadow rays of those hits,
                  50k shoot rays from 64 origins with uniform directions and
                  a max t of 80 m, 262,144 debris segments of length
                  U[1, 200] voxels
"""
from __future__ import annotations

import numpy as np

from ._native import HIT_DTYPE, RAY_DTYPE


def shadow_queries(scene, prim_rays: np.ndarray, prim_hits: np.ndarray, u: np.ndarray, lights: np.ndarray,
                   n_max, codes: bool = False):
    """Shadow rays of shade() (raymarch.wgsl:323-356) for the primary hits
    that cast them (hit, non-emissive), in pixel-then-slot order, repeated or
    cut to exactly n_max (None: every cast ray once).  Origin P + N * 0.5
    (P = ro + rd * t), sun direction with max t 1e9 when dot(N, sun) > 0; per
    light the unit vector to it with max t d - 2 when d is within its radius
    and dot(N, L) > 0.  codes=True also returns each ray's pixel and slot."""
    ok = (prim_hits["face"] != 0xFF) & ~(scene.mats.reshape(-1, 20)[prim_hits["mat"], 3] > 0)
    pix = np.nonzero(ok)[0]
    r, h = prim_rays[ok], prim_hits[ok]
    n = np.zeros((r.shape[0], 3), np.float32)
    f = h["face"].astype(np.int64)
    sel = f < 6
    n[np.nonzero(sel)[0], f[sel] >> 1] = np.where(f[sel] & 1, -1.0, 1.0)
    p = r["ro"] + r["rd"] * h["t"][:, None]
    o = (p + n * np.float32(0.5)).astype(np.float32)
    vs, sun = u[3], u[4:7]
    cand = []
    c = (n @ sun) > 0
    cand.append((c, np.broadcast_to(sun, n.shape), np.full(n.shape[0], 1e9, np.float32)))
    l4 = lights.reshape(-1, 4)
    for i in range(int(l4[0, 0])):
        lp = l4[1 + 2 * i]
        d = lp[:3] / vs - p
        dist = np.sqrt((d * d).sum(1))
        dn = d / np.maximum(dist, 1e-12)[:, None]
        cand.append(((dist <= lp[3] / vs) & ((n * dn).sum(1) > 0), dn, dist - 2.0))
    sel = np.stack([cc[0] for cc in cand], 1)  # (hits, slots)
    rows, slots = np.nonzero(sel)
    if rows.size == 0:
        raise IndexError("no primary hit casts a shadow ray")
    if n_max is not None:
        rows, slots = np.resize(rows, n_max), np.resize(slots, n_max)
    out = np.zeros(rows.shape[0], RAY_DTYPE)
    out["ro"] = o[rows]
    dirs = np.stack([cc[1] for cc in cand], 1)
    mts = np.stack([cc[2] for cc in cand], 1)
    out["rd"] = dirs[rows, slots]
    out["max_t"] = mts[rows, slots]
    out["flags"] = 1
    if codes:
        return out, pix[rows], slots
    return out


def frame_order(pixels: np.ndarray, slots: np.ndarray) -> np.ndarray:
    """The order in which the frame's shadow stage takes the rays: the queue
    region of 64 pixels (one wavefront of the primary launch), then the slot,
    then the pixel (k_frame_primary's per-region, slot-major queue)."""
    return np.lexsort((pixels, slots, pixels // 64))


def region_wave_cost(steps_in_order: np.ndarray, regions_in_order: np.ndarray, wave: int = 64) -> int:
    """Issue cost of the frame's shadow stage for rays in queue order: each
    region's rays are taken by one wavefront in chunks of 64, and a chunk
    costs its longest ray's steps."""
    s = np.asarray(steps_in_order, np.int64)
    reg = np.asarray(regions_in_order, np.int64)
    start = np.r_[0, np.nonzero(np.diff(reg))[0] + 1]
    pos = np.arange(s.size) - np.repeat(start, np.diff(np.r_[start, s.size]))
    chunk = np.cumsum(np.r_[True, (np.diff(reg) != 0) | (pos[1:] % wave == 0)]) - 1
    return int(np.maximum.reduceat(s, np.r_[0, np.nonzero(np.diff(chunk))[0] + 1]).sum())


def mixed_queries(scene, sim, gw, seed: int, vs: float = 1.0 / 160.0, width: int = 2048, height: int = 1024,
                  n_shadow: int = 3 * (1 << 20), n_shoot: int = 50000, n_seg: int = 262144):
    """The uneven query mix of the batched launch, as host arrays by kind
    (0 primary, 1 shadow, 2 shoot, 3 segment).  The shadow rays need the
    primary hits, which the plain per-kind launch provides."""
    import torch
    rng = np.random.default_rng(seed + 100)
    prim = scene.camera_rays(width, height, 90.0, seed=seed + 1, tile=8)
    u, lights = scene.frame_inputs(width, height, seed=seed + 1, voxel_size=vs, light_radius_vox=8192.0)
    d_prim = torch.from_numpy(prim.view(np.uint8).copy()).cuda()
    d_out = torch.empty(prim.shape[0] * 16, dtype=torch.uint8, device=d_prim.device)
    sim.trace_device(0, d_prim, prim.shape[0], d_out)
    gw.sync()
    ph = d_out.cpu().numpy().view(HIT_DTYPE)
    del d_prim, d_out
    shadow = shadow_queries(scene, prim, ph, u, lights, n_shadow)
    shoot = np.zeros(n_shoot, RAY_DTYPE)
    origins = np.zeros((64, 3), np.float32)
    for i in range(64):
        origins[i] = scene.camera(seed=seed + 200 + i)[:3]
    shoot["ro"] = origins[rng.integers(0, 64, n_shoot)]
    d = rng.normal(size=(n_shoot, 3))
    shoot["rd"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    shoot["max_t"] = 80.0 / vs
    shoot["flags"] = 2
    seg = np.zeros(n_seg, RAY_DTYPE)
    seg["ro"] = rng.uniform([0, 0, 0], scene.dims, size=(n_seg, 3))
    d = rng.normal(size=(n_seg, 3))
    seg["rd"] = d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(1, 200, size=(n_seg, 1))
    seg["flags"] = 3
    return {0: prim, 1: shadow, 2: shoot, 3: seg}
