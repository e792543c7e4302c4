"""GPU tests of the learned ray-cost binning (px_mlp.hip) against the CPU
oracle (oracle/mlp.py):

* forward predictions and batch gradients of the BF16/MFMA kernels equal
  the oracle's BF16-rounding restatement to FP32-accumulation-order noise;
* one Adam step on the GPU equals the oracle's Adam on the GPU gradient;
* features are in [-1, 1]; the occupancy samples are exactly +-1 and agree
  with the grid as the host reads it;
* binning is a stable permutation sorted by bucket inside every window, and
  the permuted launch gives the plain launch's hits bit for bit;
* the fused prediction + global sort equals the separate kernels' order,
  and with (direction face, bucket) keys numpy's stable sort of them;
* training on 48 cameras cuts the predictor's held-out error to under a
  quarter of the untrained network's, with a primary MAE below the mean step
  count, and the learned shadow order has a lower wavefront cost than the
  pixel-then-slot order and than the direction sort alone (the frame's own
  queue order, which beats it, is printed beside);
* the batched launch in per-kind orders (px_trace_batch_perm) gives
  px_trace_batch's hits bit for bit, with random orders and with the
  predictor's.
"""
import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(small_scene):
    from pixels_amd import GpuWorld, Simulation
    gw = GpuWorld(small_scene)
    yield small_scene, gw, Simulation(gw)
    gw.close()


def _rand_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp.N_PARAMS, np.float32)
    p[mlp.OFF_W1:mlp.OFF_B1] = rng.uniform(-0.6, 0.6, mlp.N_IN * mlp.N_H)
    p[mlp.OFF_B1:mlp.OFF_W2] = rng.uniform(-0.1, 0.3, mlp.N_H)
    p[mlp.OFF_W2:mlp.OFF_B2] = rng.uniform(-0.25, 0.25, mlp.N_H * mlp.N_H)
    p[mlp.OFF_B2:mlp.OFF_W3] = rng.uniform(-0.1, 0.3, mlp.N_H)
    p[mlp.OFF_W3:mlp.OFF_B3] = rng.uniform(-0.3, 0.3, mlp.N_H)
    p[mlp.OFF_B3] = 1.5
    return p


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_forward_and_gradient_match_oracle(world):
    import torch
    from pixels_amd.learned import CostModel
    s, gw, sim = world
    m = CostModel(gw, seed=7)
    try:
        p = _rand_params(11)
        m.set_params(p)
        assert np.array_equal(m.params(), p)
        rng = np.random.default_rng(12)
        n = 1000  # not a multiple of 64
        x = mlp.bf16(rng.uniform(-1, 1, (n, 16)).astype(np.float32))
        xb = (x.view(np.uint32) >> 16).astype(np.uint16)
        t = rng.uniform(0, 8, n).astype(np.float32)
        dx, dt = _dev(xb), _dev(t)
        pred = torch.zeros(n, dtype=torch.float32, device="cuda")
        print("case: mlp forward n=1000", flush=True)
        m.predict(dx, n, pred)
        gw.sync()
        got = pred.cpu().numpy()
        ref, _, _ = mlp.forward_bf16(p, x)
        err = np.abs(got - ref)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 5e-3, err.max()
        assert np.quantile(err, 0.99) < 0.05
        print("case: mlp gradient n=1000", flush=True)
        loss, g = m.grad(dx, dt, n)
        rloss, rg = mlp.grad_bf16(p, x, t)
        assert abs(loss - rloss) <= 2e-3 * rloss
        assert np.linalg.norm(g - rg) / np.linalg.norm(rg) < 2e-2
        for lo, hi in ((mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_W2, mlp.OFF_B2), (mlp.OFF_W3, mlp.OFF_B3)):
            gr, rr = g[lo:hi], rg[lo:hi]
            cos = float(np.dot(gr, rr) / (np.linalg.norm(gr) * np.linalg.norm(rr)))
            assert cos > 0.999, (lo, cos)
        # one Adam step equals the oracle's Adam on the kernel's own gradient
        print("case: adam step", flush=True)
        m.set_params(p)
        _, g0 = m.grad(dx, dt, n)
        m.train_step(dx, dt, n, lr=1e-3)
        p1 = m.params()
        z = np.zeros_like(p)
        rp1, _, _ = mlp.adam(p, g0, z, z, 1, lr=1e-3)
        assert np.allclose(p1, rp1, rtol=1e-6, atol=1e-7)
    finally:
        m.close()


def test_features_range(world):
    import torch
    from pixels_amd.learned import CostModel
    s, gw, sim = world
    m = CostModel(gw)
    try:
        rays = s.camera_rays(320, 180, 90.0, seed=4)
        n = rays.shape[0]
        feat = torch.zeros(n * 16, dtype=torch.int16, device="cuda")
        m.features(0, _dev(rays.view(np.uint8)), n, feat)
        gw.sync()
        f = mlp.bf16_bits_to_f32(feat.cpu().numpy().view(np.uint16)).reshape(n, 16)
        assert np.all(f >= -1.0) and np.all(f <= 1.0)
        occ = np.c_[f[:, 6:15], f[:, 1]]  # samples at 0.5, 4, .., 512 voxels, then at 2
        assert np.all(np.isin(occ, (-1.0, 1.0)))
        assert np.all(f[:, 15] == -1.0)
        d = rays["rd"] / np.linalg.norm(rays["rd"], axis=1, keepdims=True)
        assert np.allclose(f[:, 3:6], d, atol=1e-2)
        # the occupancy samples, recomputed on the host from the grid: a
        # sample is +1 when the brick (or super cell) at that distance past
        # the entry is non-empty; rounding of the hardware rcp / rsq moves a
        # few points across a cell face
        dims = np.array(s.dims, np.float64)
        bd, sd = dims.astype(np.int64) // 8, (dims.astype(np.int64) // 8 + 3) // 4
        grid = s.grid.reshape(bd[2], bd[1], bd[0]) != 0
        sup = np.zeros(sd[::-1], bool)
        zz, yy, xx = np.nonzero(grid)
        sup[zz // 4, yy // 4, xx // 4] = True
        ro, rd = rays["ro"].astype(np.float64), rays["rd"].astype(np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            t0, t1 = -ro / rd, (dims - ro) / rd
        tmn = np.maximum(np.nanmax(np.minimum(t0, t1), 1), 0.0)
        tmx = np.minimum(np.nanmin(np.maximum(t0, t1), 1), rays["max_t"])
        ln = np.linalg.norm(rd, axis=1)
        ext = np.maximum(np.minimum(tmx, tmn + dims.max()) - tmn, 0.0) * ln
        want = np.empty((n, 10))
        for k, dist in enumerate((0.5, 4, 8, 16, 32, 64, 128, 256, 512, 2)):
            cell = 8.0 if (k < 5 or k == 9) else 32.0
            p = ro + rd * (tmn + dist / ln)[:, None]
            c = np.floor(p / cell).astype(np.int64)
            lim = bd if cell == 8.0 else sd
            inb = (ext > 0) & (dist <= ext) & (c >= 0).all(1) & (c < lim).all(1)
            cc = np.where(inb[:, None], c, 0)
            o = (grid if cell == 8.0 else sup)[cc[:, 2], cc[:, 1], cc[:, 0]] & inb
            want[:, k] = np.where(o, 1.0, -1.0)
        agree = (want == occ).mean(0)
        assert agree.min() > 0.995, agree
        # rays inside occupied space see some occupied cells
        assert (occ > 0).mean() > 0.2
    finally:
        m.close()


def _trace_steps(sim, gw, kind, rays):
    import torch
    n = rays.shape[0]
    out = torch.zeros(n * 16, dtype=torch.uint8, device="cuda")
    sim.trace_device(kind, _dev(rays.view(np.uint8)), n, out)
    gw.sync()
    return out.cpu().numpy().view(mlp_hit_dtype())


def mlp_hit_dtype():
    import oracle
    return oracle.HIT_DTYPE


def test_binning_is_stable_permutation_and_bit_exact(world):
    import torch
    from pixels_amd.learned import bin_rays, trace_perm
    s, gw, sim = world
    rays = s.camera_rays(480, 270, 90.0, seed=8)
    n = rays.shape[0]
    rng = np.random.default_rng(3)
    pred = rng.uniform(-1, 10, n).astype(np.float32)
    perm = torch.zeros(n, dtype=torch.int32, device="cuda")
    win, nb, lo, hi = 1000, 32, 0.0, 9.0
    print("case: bin n=%d" % n, flush=True)
    bin_rays(gw, _dev(pred), n, perm, window=win, buckets=nb, lo=lo, hi=hi)
    gw.sync()
    pm = perm.cpu().numpy().astype(np.int64)
    b = np.clip(np.floor((pred - lo) * (nb / (hi - lo))), 0, nb - 1).astype(np.int64)
    for w0 in range(0, n, win):
        seg = pm[w0:w0 + win]
        assert np.array_equal(np.sort(seg), np.arange(w0, min(w0 + win, n)))
        exp = w0 + np.lexsort((np.arange(seg.size), b[w0:w0 + win]))
        assert np.array_equal(seg, exp)
    for kind in (0, 1):
        ref = _trace_steps(sim, gw, kind, rays)
        out = torch.zeros(n * 16, dtype=torch.uint8, device="cuda")
        trace_perm(gw, kind, _dev(rays.view(np.uint8)), perm, n, out)
        gw.sync()
        assert np.array_equal(out.cpu().numpy(), ref.view(np.uint8))


@pytest.mark.parametrize("kind", [0, 1])
def test_fused_global_sort_equals_separate_kernels(world, kind):
    """px_mlp_sort_rays (features + network + global counting sort, three
    launches) gives exactly the permutation of features -> predict ->
    bin_rays(window = n), which is numpy's stable sort of the buckets; the
    permuted launch then equals the plain launch.  Ragged counts (not whole
    256-ray tiles, fewer rays than one tile, one ray) and a smaller second
    call on the same scratch are covered."""
    import torch
    from pixels_amd.learned import CostModel, bin_rays, trace_perm
    s, gw, sim = world
    m = CostModel(gw, seed=5)
    try:
        m.set_params(_rand_params(21 + kind))
        rays_all = s.camera_rays(480, 270, 90.0, seed=9 + kind)
        if kind == 1:  # shadow-like rays: from the primary hits toward a fixed sun, bounded
            hits = _trace_steps(sim, gw, 0, rays_all)
            ok = hits["face"] != 0xFF
            r = rays_all[ok].copy()
            r["ro"] = r["ro"] + r["rd"] * (hits["t"][ok, None] - 0.5)
            sun = np.array([0.3, 0.8, 0.2], np.float32)
            r["rd"] = sun / np.linalg.norm(sun)
            r["max_t"] = 1e9
            rays_all = r
        nb = 32
        for n in (rays_all.shape[0], 1000, 255, 1):
            rays = rays_all[:n]
            d = _dev(rays.view(np.uint8))
            feat = torch.zeros(n * 16, dtype=torch.int16, device="cuda")
            pred = torch.zeros(n, dtype=torch.float32, device="cuda")
            m.features(kind, d, n, feat)
            m.predict(feat, n, pred)
            gw.sync()
            pr = pred.cpu().numpy()
            lo, hi = float(np.quantile(pr, 0.02)), float(np.quantile(pr, 0.98)) + 1e-3
            if n == 1:
                lo, hi = float(pr[0]) - 1.0, float(pr[0]) + 1.0
            ref = torch.zeros(n, dtype=torch.int32, device="cuda")
            bin_rays(gw, pred, n, ref, window=n, buckets=nb, lo=lo, hi=hi)
            perm = torch.full((n,), -1, dtype=torch.int32, device="cuda")
            print("case: fused sort kind=%d n=%d" % (kind, n), flush=True)
            m.sort_rays(kind, d, n, perm, buckets=nb, lo=lo, hi=hi)
            gw.sync()
            pm, rf = perm.cpu().numpy().astype(np.int64), ref.cpu().numpy().astype(np.int64)
            scale = np.float32(nb) / (np.float32(hi) - np.float32(lo))
            b = np.clip(np.floor((pr - np.float32(lo)) * scale), 0, nb - 1).astype(np.int64)
            assert np.array_equal(pm, rf)
            assert np.array_equal(pm, np.lexsort((np.arange(n), b)))
            # (direction face, bucket) keys: 6 * nb = 192 keys
            rd = rays["rd"].astype(np.float32)
            ad = np.abs(rd)
            ax = np.where(ad[:, 1] > ad[:, 0], 1, 0)
            ax = np.where(ad[:, 2] > ad[np.arange(n), ax], 2, ax)
            face = 2 * ax + (rd[np.arange(n), ax] < 0)
            permf = torch.full((n,), -1, dtype=torch.int32, device="cuda")
            m.sort_rays(kind, d, n, permf, buckets=nb, lo=lo, hi=hi, faces=True)
            gw.sync()
            assert np.array_equal(permf.cpu().numpy().astype(np.int64), np.lexsort((np.arange(n), b, face)))
            # windows of 2^k consecutive rays, each sorted on its own (k = 17
            # holds all 129,600 rays: the global sort again)
            for wlog in (8, 12, 17):
                win = np.arange(n) >> wlog
                print("case: windowed sort kind=%d n=%d window=2^%d" % (kind, n, wlog), flush=True)
                for faces, exp in ((True, np.lexsort((np.arange(n), b, face, win))),
                                   (False, np.lexsort((np.arange(n), b, win)))):
                    permw = torch.full((n,), -1, dtype=torch.int32, device="cuda")
                    m.sort_rays(kind, d, n, permw, buckets=nb, lo=lo, hi=hi, faces=faces, window_log2=wlog)
                    gw.sync()
                    assert np.array_equal(permw.cpu().numpy().astype(np.int64), exp), (faces, wlog)
            if n > 1000:
                assert len(np.unique(b)) > nb // 2  # the buckets are really used
                plain = _trace_steps(sim, gw, kind, rays)
                out = torch.zeros(n * 16, dtype=torch.uint8, device="cuda")
                trace_perm(gw, kind, d, perm, n, out)
                gw.sync()
                assert np.array_equal(out.cpu().numpy(), plain.view(np.uint8))
                # the one-call form: same order, same hits
                out2 = torch.zeros(n * 16, dtype=torch.uint8, device="cuda")
                perm2 = torch.full((n,), -1, dtype=torch.int32, device="cuda")
                m.trace_sorted(kind, d, n, perm2, out2, buckets=nb, lo=lo, hi=hi)
                gw.sync()
                assert torch.equal(perm2, perm) and np.array_equal(out2.cpu().numpy(), plain.view(np.uint8))
                m.trace_sorted(kind, d, n, perm2, out2, buckets=nb, lo=lo, hi=hi, faces=True)
                gw.sync()
                assert torch.equal(perm2, permf) and np.array_equal(out2.cpu().numpy(), plain.view(np.uint8))
    finally:
        m.close()


def test_sort_rays_rejects_bad_arguments(world):
    """Unknown flags, more than 256 keys (42 buckets with faces; a bucket
    count whose 6x overflows 32 bits), an empty bucket range, a window below
    one 256-ray tile or above 2^30 rays and an unknown kind are errors, and
    nothing is written."""
    import torch
    from pixels_amd import PxError
    from pixels_amd.learned import CostModel
    s, gw, sim = world
    m = CostModel(gw, seed=2)
    try:
        rays = s.camera_rays(64, 32, 90.0, seed=3)
        n = rays.shape[0]
        d = _dev(rays.view(np.uint8))
        perm = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        bad = [dict(buckets=43, faces=True), dict(buckets=257), dict(buckets=0), dict(buckets=0x2AAAAAAB, faces=True),
               dict(lo=3.0, hi=3.0), dict(window_log2=7), dict(window_log2=31)]
        for kw in bad:
            with pytest.raises(PxError):
                m.sort_rays(1, d, n, perm, **kw)
        with pytest.raises(PxError):
            m.sort_rays(7, d, n, perm)
        with pytest.raises(PxError):
            gw.check(gw.lib.px_mlp_sort_rays(m.h, 1, d.data_ptr(), n, 32, 0.0, 9.0, 2, perm.data_ptr()))
        gw.sync()
        assert bool((perm == -1).all())
        m.sort_rays(1, d, n, perm, buckets=42, faces=True)  # the largest face-keyed sort
        gw.sync()
        assert np.array_equal(np.sort(perm.cpu().numpy()), np.arange(n))
    finally:
        m.close()


def test_training_gain_over_untrained(world):
    """Training on the DDA's own step counts over 48 seeded 160x90 cameras
    makes the predictor carry over to a held-out 640x360 camera: its held-out
    log2 error is under a quarter of the untrained network's and its MAE in
    steps is below the mean step count.  On that camera's shadow rays the
    learned order (direction face, then predicted cost) has a lower
    wavefront cost (the sum over 64-ray wavefronts of the longest ray's
    steps, deterministic) than the pixel-then-slot order and than the
    direction sort alone; the frame's own queue order is printed beside it
    (per-region, slot-major; it beats the learned order, as the measured
    frame shows)."""
    import torch
    from pixels_amd.learned import CostModel, log2_steps, wave_cost
    from pixels_amd.workloads import frame_order, region_wave_cost, shadow_queries
    s, gw, sim = world

    def data(seed, w, h):
        prim = s.camera_rays(w, h, 90.0, seed=seed, tile=8)
        ph = _trace_steps(sim, gw, 0, prim)
        u, lights = s.frame_inputs(w, h, seed=seed)
        sh, pix, slot = shadow_queries(s, prim, ph, u, lights, None, codes=True)
        hs = _trace_steps(sim, gw, 1, sh)
        return {0: (prim, ph["steps"].astype(np.int64)), 1: (sh, hs["steps"].astype(np.int64))}, (pix, slot)

    def feats(m, kind, rays):
        n = rays.shape[0]
        f = torch.zeros(n * 16, dtype=torch.int16, device="cuda")
        m.features(kind, _dev(rays.view(np.uint8)), n, f)
        return f

    models = {k: CostModel(gw, seed=3 + k) for k in (0, 1)}
    try:
        untrained = {k: m.params() for k, m in models.items()}
        train = {0: ([], []), 1: ([], [])}
        cams = 0
        for cam in range(200, 400):
            try:
                d, _ = data(cam, 160, 90)
            except IndexError:  # no shadow rays from this camera
                continue
            for k in (0, 1):
                train[k][0].append(feats(models[k], k, d[k][0]))
                train[k][1].append(_dev(log2_steps(d[k][1])))
            cams += 1
            if cams == 48:
                break
        held, (pix, slot) = data(47, 640, 360)
        bs = 65536
        for k in (0, 1):
            tf, tt = torch.cat(train[k][0]), torch.cat(train[k][1])
            n = tt.numel()
            order = torch.randperm(n, device="cuda", generator=torch.Generator(device="cuda").manual_seed(k))
            tf = tf.view(n, 16)[order].reshape(-1).contiguous()
            tt = tt[order].contiguous()
            print("case: training kind %d on %d rays from %d cameras" % (k, n, cams), flush=True)
            for ep in range(12):
                for off in range(0, n - bs + 1, bs):
                    models[k].train_step(tf[off * 16:], tt[off:], bs, lr=1e-3)
            # the steps run on the library's stream: wait before torch may
            # hand tf / tt's memory to the next tensors
            gw.sync()
        report = {}
        for k in (0, 1):
            m = models[k]
            rays, steps = held[k]
            nh = rays.shape[0]
            f = feats(m, k, rays)
            target = _dev(log2_steps(steps))
            nb = nh // 64 * 64
            trained = m.params()
            losses = {}
            for name, params in (("untrained", untrained[k]), ("trained", trained)):
                m.set_params(params)
                losses[name], _ = m.grad(f, target, nb)
            pred = torch.zeros(nh, dtype=torch.float32, device="cuda")
            m.predict(f, nh, pred)
            gw.sync()
            mae = float(np.mean(np.abs(np.exp2(pred.cpu().numpy()) - np.maximum(steps, 1))))
            report[k] = (losses, mae, float(steps.mean()))
            assert losses["trained"] < 0.25 * losses["untrained"], (k, losses)
            if k == 0:
                assert mae < steps.mean(), (mae, steps.mean())
        # wavefront cost of the held-out shadow rays in each order
        sh, steps = held[1]
        nh = sh.shape[0]
        d = _dev(sh.view(np.uint8))
        perm = torch.zeros(nh, dtype=torch.int32, device="cuda")
        models[1].sort_rays(1, d, nh, perm, buckets=32, faces=True)
        gw.sync()
        c_learned = wave_cost(steps[perm.cpu().numpy().astype(np.int64)])
        models[1].sort_rays(1, d, nh, perm, buckets=1, faces=True)
        gw.sync()
        c_faces = wave_cost(steps[perm.cpu().numpy().astype(np.int64)])
        c_plain = wave_cost(steps)
        fo = frame_order(pix, slot)
        c_frame = region_wave_cost(steps[fo], (pix // 64)[fo])
        print("held-out: primary log2 MSE %.3f -> %.3f, MAE %.2f steps (mean %.1f); shadow log2 MSE %.3f -> %.3f, "
              "MAE %.2f (mean %.1f); shadow wave cost: pixel-then-slot %d, faces only %d, learned %d, frame queue "
              "order %d" % (report[0][0]["untrained"], report[0][0]["trained"], report[0][1], report[0][2],
                            report[1][0]["untrained"], report[1][0]["trained"], report[1][1], report[1][2], c_plain,
                            c_faces, c_learned, c_frame), flush=True)
        assert c_learned < 0.9 * c_plain
        assert c_learned < c_faces
    finally:
        for m in models.values():
            m.close()


def test_batch_in_learned_order_equals_batch(world):
    import torch
    from pixels_amd.learned import CostModel
    from pixels_amd.workloads import shadow_queries
    s, gw, sim = world
    prim = s.camera_rays(320, 180, 90.0, seed=5, tile=8)
    ph = _trace_steps(sim, gw, 0, prim)
    u, lights = s.frame_inputs(320, 180, seed=5)
    sh = shadow_queries(s, prim, ph, u, lights, None)
    rng = np.random.default_rng(6)
    seg = np.zeros(5000, prim.dtype)
    seg["ro"] = rng.uniform((0, 0, 0), s.dims, (5000, 3))
    seg["rd"] = rng.normal(size=(5000, 3)) * 20
    seg["flags"] = 3
    host = {0: prim, 1: sh, 3: seg}
    rays = {k: (_dev(v.view(np.uint8)), v.shape[0]) for k, v in host.items()}
    outs = {k: torch.full((v.shape[0] * 16,), 0xAB, dtype=torch.uint8, device="cuda") for k, v in host.items()}
    q = sim.query_set(rays, outs)
    sim.trace_batch(q)
    gw.sync()
    ref = {k: o.clone() for k, o in outs.items()}
    m = CostModel(gw, seed=9)
    try:
        for how in ("random", "learned"):
            perms = {}
            for k in (0, 1, 3):
                n = host[k].shape[0]
                if how == "random" or k == 3:
                    perms[k] = torch.from_numpy(rng.permutation(n).astype(np.int32)).cuda()
                else:
                    perms[k] = torch.zeros(n, dtype=torch.int32, device="cuda")
                    m.sort_rays(k, rays[k][0], n, perms[k], buckets=32, faces=True)
            for o in outs.values():
                o.fill_(0xCD)
            print("case: batch in %s order" % how, flush=True)
            sim.trace_batch_perm(q, perms)
            gw.sync()
            for k in host:
                assert torch.equal(outs[k], ref[k]), (how, k)
    finally:
        m.close()
