"""GPU parity at the measured workloads' full sizes, bit-exact against the
oracle:

* the headline frame: the 1280x352x960 scene (2000 primitives, structural
  floor and ceiling slabs, 10% glass), 1920x1080 primary rays in 8x8 tiles
  plus the sun and 8 point-light shadow rays of shade()
  (raymarch.wgsl:323-356, 388-398); every primary and every shadow slot is
  compared, and a ragged ray count covers the partial last region (which
  the shadow kernel issues first);
* the batched mix on the 2560x704x1920 scene (generator scaled x2, 160
  voxels per metre): 2M primary, 3M shadow, 50k shoot rays from 64 origins,
  262,144 debris segments through one graph launch of px_trace_batch, equal
  byte for byte to the four per-kind launches, plus an oracle sample of
  every kind, and again with the shadow queries in a learned order
  (px_mlp_sort_rays, then px_trace_batch_perm);
* learned binning on the 512x256x512 scene with the structural floor: the
  binned launch (px_mlp_trace_sorted) gives the plain launch's hits bit for
  bit on 2M shadow rays.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _mismatch(a, b):
    a = a.view(np.uint8).reshape(a.shape[0], -1)
    b = b.view(np.uint8).reshape(b.shape[0], -1)
    return np.nonzero((a != b).any(1))[0]


def _frame(sim, gw, rays, n, slots):
    import torch
    dr = torch.from_numpy(rays[:n].view(np.uint8).copy()).cuda()
    dp = torch.full((n * 16,), 0xAB, dtype=torch.uint8, device="cuda")
    ds = torch.full((n * slots * 16,), 0xAB, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    nc = sim.trace_frame_device(dr, n, dp, ds, count=True)
    gw.sync()
    return dp.cpu().numpy().view(oracle.HIT_DTYPE), ds.cpu().numpy().view(oracle.HIT_DTYPE), nc


@pytest.fixture(scope="module")
def headline():
    from pixels_amd import GpuWorld, Scene, Simulation
    from pixels_amd.scene import LARGE
    print("case: headline scene 1280x352x960 seed 2", flush=True)
    s = Scene.build(LARGE, seed=2)
    assert s.dims == (1280, 352, 960)
    gw = GpuWorld(s)
    sim = Simulation(gw)
    rays = s.camera_rays(1920, 1080, 90.0, seed=2, tile=8)
    u, lights = s.frame_inputs(1920, 1080, seed=2)
    gw.set_uniforms(u)
    gw.set_lights(lights)
    yield s, gw, sim, oracle.World.from_scene(s), rays, u, lights
    gw.close()


@pytest.mark.parametrize("n,planar", [(1920 * 1080, False), (1920 * 1080 - 37, False), (1920 * 1080 - 37, True)])
def test_headline_frame_bit_exact(headline, n, planar):
    s, gw, sim, ow, rays, u, lights = headline
    slots = sim.frame_slots()
    assert slots == 9
    print("case: headline frame n=%d slot-planar=%s" % (n, planar), flush=True)
    sim.set_frame_layout(planar)
    try:
        prim, sh, nc = _frame(sim, gw, rays, n, slots)
    finally:
        sim.set_frame_layout(False)
    if planar:  # slot s of ray i at s * n + i -> pixel-major
        sh = np.ascontiguousarray(sh.reshape(slots, n).T).reshape(-1)
    rp, rs, rnc = ow.frame(u, lights, rays[:n])
    bad = _mismatch(prim, rp)
    assert bad.size == 0, "primary: %d mismatches, first %s" % (bad.size, bad[:5])
    bad = _mismatch(sh, rs)
    assert bad.size == 0, "shadow: %d mismatches, first %s" % (bad.size, bad[:5])
    assert nc == rnc
    # the scene and camera of the measured frame: most rays hit, ~6 shadow rays per pixel
    assert (rp["face"] != 0xFF).mean() > 0.5
    assert nc > 4 * n


def test_batched_mix_full_size():
    import torch
    from pixels_amd import GpuWorld, Scene, Simulation
    from pixels_amd.scene import LARGE_SCALED
    from pixels_amd.workloads import mixed_queries
    print("case: batched scene 2560x704x1920 seed 2", flush=True)
    s = Scene.build(LARGE_SCALED, seed=2, dims=(2560, 704, 1920), scale=2.0)
    gw = GpuWorld(s, voxel_size=1.0 / 160.0)
    try:
        sim = Simulation(gw)
        host = mixed_queries(s, sim, gw, 2)
        counts = {k: v.shape[0] for k, v in host.items()}
        assert counts == {0: 2048 * 1024, 1: 3 << 20, 2: 50000, 3: 262144}
        rays = {k: torch.from_numpy(v.view(np.uint8).copy()).cuda() for k, v in host.items()}
        outb = {k: torch.full((v.shape[0] * 16,), 0xAB, dtype=torch.uint8, device="cuda") for k, v in host.items()}
        outp = {k: torch.full((v.shape[0] * 16,), 0xCD, dtype=torch.uint8, device="cuda") for k, v in host.items()}
        q = sim.query_set({k: (rays[k], counts[k]) for k in host}, outb)
        torch.cuda.synchronize()
        print("case: batched graph launch, counts %s" % counts, flush=True)
        sim.trace_batch(q)
        for k in host:
            sim.trace_device(k, rays[k], counts[k], outp[k])
        gw.sync()
        ow = oracle.World.from_scene(s)
        for k in host:
            got = outb[k].cpu().numpy()
            assert np.array_equal(got, outp[k].cpu().numpy()), "kind %d: batch differs from per-kind launch" % k
            hits = got.view(oracle.HIT_DTYPE)
            idx = np.linspace(0, counts[k] - 1, 3000).astype(np.int64)
            exp = ow.trace(k, host[k][idx])
            bad = _mismatch(hits[idx], exp)
            assert bad.size == 0, "kind %d: %d oracle mismatches in the sample" % (k, bad.size)
        # a replay of the captured graph gives the same bytes again
        for t in outb.values():
            t.fill_(0x11)
        sim.trace_batch(q)
        gw.sync()
        for k in host:
            assert np.array_equal(outb[k].cpu().numpy(), outp[k].cpu().numpy()), "replay, kind %d" % k
        # the learned-binned form: the 3M shadow queries in the predictor's
        # (direction face, cost bucket) order inside windows of 2^18 queries,
        # through px_trace_batch_perm
        from pixels_amd.learned import CostModel
        m = CostModel(gw, seed=4)
        try:
            perm = torch.zeros(counts[1], dtype=torch.int32, device="cuda")
            m.sort_rays(1, rays[1], counts[1], perm, buckets=32, faces=True, window_log2=18)
            for t in outb.values():
                t.fill_(0x22)
            print("case: batched launch, shadow kind in the learned order", flush=True)
            sim.trace_batch_perm(q, {1: perm})
            gw.sync()
            pm = perm.cpu().numpy()
            assert np.array_equal(np.sort(pm), np.arange(counts[1], dtype=np.int32)), "not a permutation"
            for k in host:
                assert np.array_equal(outb[k].cpu().numpy(), outp[k].cpu().numpy()), "learned order, kind %d" % k
        finally:
            m.close()
    finally:
        gw.close()


def test_learned_binning_floor_scene_bit_identical(floor_scene):
    import torch
    from pixels_amd import GpuWorld, Simulation
    from pixels_amd.learned import CostModel, log2_steps
    from pixels_amd.workloads import shadow_queries
    s = floor_scene
    gw = GpuWorld(s)
    try:
        sim = Simulation(gw)
        W, H = 1920, 1080
        prim = s.camera_rays(W, H, 90.0, seed=309, tile=8)
        d_prim = torch.from_numpy(prim.view(np.uint8).copy()).cuda()
        out = torch.empty(prim.shape[0] * 16, dtype=torch.uint8, device="cuda")
        sim.trace_device(0, d_prim, prim.shape[0], out)
        gw.sync()
        ph = out.cpu().numpy().view(oracle.HIT_DTYPE)
        u, lights = s.frame_inputs(W, H, seed=309)
        sh = shadow_queries(s, prim, ph, u, lights, W * H)
        n = sh.shape[0]
        d = torch.from_numpy(sh.view(np.uint8).copy()).cuda()
        plain = torch.empty(n * 16, dtype=torch.uint8, device="cuda")
        sim.trace_device(1, d, n, plain)
        m = CostModel(gw, seed=5)
        try:
            # a short training on the plain launch's own step counts
            gw.sync()
            steps = plain.cpu().numpy().view(oracle.HIT_DTYPE)["steps"]
            f = torch.empty(n * 16, dtype=torch.int16, device="cuda")
            m.features(1, d, n, f)
            t = torch.from_numpy(log2_steps(steps)).cuda()
            bs = 65536
            for off in range(0, n - bs + 1, bs):
                m.train_step(f[off * 16:], t[off:], bs, lr=1e-3)
            perm = torch.full((n,), -1, dtype=torch.int32, device="cuda")
            binned = torch.full((n * 16,), 0xAB, dtype=torch.uint8, device="cuda")
            print("case: learned binning on the floor scene, %d shadow rays" % n, flush=True)
            m.trace_sorted(1, d, n, perm, binned, faces=True)
            gw.sync()
            assert np.array_equal(np.sort(perm.cpu().numpy()), np.arange(n))
            assert torch.equal(plain, binned)
            ow = oracle.World.from_scene(s)
            idx = np.linspace(0, n - 1, 4000).astype(np.int64)
            bad = _mismatch(binned.cpu().numpy().view(oracle.HIT_DTYPE)[idx], ow.trace(1, sh[idx]))
            assert bad.size == 0
        finally:
            m.close()
    finally:
        gw.close()
