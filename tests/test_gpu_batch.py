"""Batched queries: all four kinds in one persistent launch replayed as a HIP
graph (px_trace_batch) give, under every launch shape, exactly the hits of
the per-kind launches (px_trace_device), which the oracle tests pin; a sample
is checked against the oracle directly.  Graph replays are checked one by one
with the program's own eager work between them, and the autotuner leaves a
shape that still gives the same bits."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _queries(s, rng, n_prim, n_shadow, n_shoot, n_seg):
    """Uneven query mix on scene s (host arrays, by kind)."""
    prim = s.camera_rays(640, 360, 90.0, seed=21)[:n_prim]
    sh = np.zeros(n_shadow, oracle.RAY_DTYPE)
    sh["ro"] = rng.uniform([0, 0, 0], s.dims, size=(n_shadow, 3))
    d = rng.normal(size=(n_shadow, 3))
    sh["rd"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    sh["max_t"] = np.where(rng.random(n_shadow) < 0.5, 1e9, rng.uniform(1, 300, n_shadow))
    sho = np.zeros(n_shoot, oracle.RAY_DTYPE)
    origins = rng.uniform([0, 0, 0], s.dims, size=(64, 3))
    sho["ro"] = origins[rng.integers(0, 64, n_shoot)]
    d = rng.normal(size=(n_shoot, 3))
    sho["rd"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    sho["max_t"] = 80.0 * 80.0
    seg = np.zeros(n_seg, oracle.RAY_DTYPE)
    seg["ro"] = rng.uniform([0, 0, 0], s.dims, size=(n_seg, 3))
    d = rng.normal(size=(n_seg, 3))
    seg["rd"] = d / np.linalg.norm(d, axis=1, keepdims=True) * rng.uniform(1, 200, size=(n_seg, 1))
    return {0: prim, 1: sh, 2: sho, 3: seg}


class _Batch:
    def __init__(self, s, q):
        import torch
        from pixels_amd import GpuWorld, Simulation
        self.gw = GpuWorld(s)
        self.sim = Simulation(self.gw)
        self.host = q
        self.rays = {k: torch.from_numpy(np.ascontiguousarray(v).view(np.uint8).copy()).cuda() for k, v in q.items()}
        self.out = {k: torch.zeros(max(1, v.shape[0]) * 16, dtype=torch.uint8, device="cuda") for k, v in q.items()}
        self.qs = self.sim.query_set({k: (self.rays[k], q[k].shape[0]) for k in q}, self.out)
        torch.cuda.synchronize()

    def results(self):
        self.gw.sync()
        return {k: self.out[k].cpu().numpy().view(oracle.HIT_DTYPE)[:self.host[k].shape[0]] for k in self.host}

    def per_kind(self):
        import torch
        ref = {}
        for k, v in self.host.items():
            o = torch.zeros(max(1, v.shape[0]) * 16, dtype=torch.uint8, device="cuda")
            if v.shape[0]:
                self.sim.trace_device(k, self.rays[k], v.shape[0], o)
            self.gw.sync()
            ref[k] = o.cpu().numpy().view(oracle.HIT_DTYPE)[:v.shape[0]]
        return ref

    def clear(self):
        for t in self.out.values():
            t.fill_(0xAB)


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.fixture(scope="module")
def batch(small_scene):
    rng = np.random.default_rng(5)
    q = _queries(small_scene, rng, 70001, 130003, 5003, 40001)
    print("case: batch on small scene, counts", {k: v.shape[0] for k, v in q.items()}, flush=True)
    b = _Batch(small_scene, q)
    yield b
    b.gw.close()


def test_per_kind_launches_match_oracle_sample(batch, small_scene):
    ow = oracle.World.from_scene(small_scene)
    ref = batch.per_kind()
    for k, v in batch.host.items():
        idx = np.arange(0, v.shape[0], 97)
        got = ref[k][idx]
        exp = ow.trace(k, v[idx])
        assert _same(got, exp), "kind %d differs from the oracle" % k


SHAPES = [(256, 0, 256), (64, 37, 1), (128, 4096, 1024), (256, 1, 7), (64, 8192, 64)]


@pytest.mark.parametrize("shape", SHAPES)
def test_batch_equals_per_kind(batch, shape):
    ref = batch.per_kind()
    batch.sim.set_batch_shape(*shape)
    print("case: batch shape", shape, flush=True)
    batch.clear()
    batch.sim.trace_batch(batch.qs)
    got = batch.results()
    for k in ref:
        assert _same(got[k], ref[k]), "kind %d differs under shape %s" % (k, shape)


def test_graph_replays_with_eager_work_between(batch):
    import torch
    ref = batch.per_kind()
    batch.sim.set_batch_shape(256, 0, 256)
    extra = torch.from_numpy(batch.host[0][:5000].view(np.uint8).copy()).cuda()
    extra_out = torch.zeros(5000 * 16, dtype=torch.uint8, device="cuda")
    for rep in range(4):
        print("case: graph replay", rep, flush=True)
        batch.clear()
        batch.sim.trace_batch(batch.qs)
        got = batch.results()
        for k in ref:
            assert _same(got[k], ref[k]), "replay %d kind %d" % (rep, k)
        # eager work of the program between replays
        batch.sim.trace_device(1, extra, 5000, extra_out)
        (extra_out.float() * 2).sum().item()
        batch.gw.sync()


def test_autotune_keeps_results(batch):
    ref = batch.per_kind()
    shape, ms = batch.sim.autotune_batch(batch.qs, repeats=1)
    print("case: autotuned", shape, ms, flush=True)
    assert shape["block"] in (64, 128, 256) and shape["chunk"] in (64, 256, 1024) and ms > 0
    assert batch.sim.get_batch_shape() == shape
    batch.clear()
    batch.sim.trace_batch(batch.qs)
    got = batch.results()
    for k in ref:
        assert _same(got[k], ref[k])
    # the shape for the launch in given orders (px_autotune_batch_perm), here
    # each kind reversed: the tuner leaves its pick set and the permuted
    # launch still gives the per-kind hits
    import torch
    perms = {k: torch.arange(int(batch.rays[k].numel()) // 32 - 1, -1, -1, dtype=torch.int32, device="cuda")
             for k in range(4) if int(batch.rays[k].numel()) >= 32}
    shape_p, ms_p = batch.sim.autotune_batch(batch.qs, repeats=1, perms=perms)
    print("case: autotuned for reversed orders", shape_p, ms_p, flush=True)
    assert shape_p["block"] in (64, 128, 256) and shape_p["chunk"] in (64, 256, 1024) and ms_p > 0
    assert batch.sim.get_batch_shape() == shape_p
    batch.clear()
    batch.sim.trace_batch_perm(batch.qs, perms)
    got = batch.results()
    for k in ref:
        assert _same(got[k], ref[k])


def test_empty_and_one_kind(small_scene):
    from pixels_amd import PxError
    rng = np.random.default_rng(6)
    q = _queries(small_scene, rng, 0, 0, 0, 777)
    b = _Batch(small_scene, q)
    try:
        ref = b.per_kind()
        b.sim.trace_batch(b.qs)
        got = b.results()
        assert _same(got[3], ref[3])
        # all empty: a no-op
        q0 = b.sim.query_set({k: (b.rays[k], 0) for k in range(4)}, b.out)
        b.sim.trace_batch(q0)
        b.gw.sync()
        # a non-empty kind without buffers is an error
        bad = b.sim.query_set({k: (b.rays[k], 0) for k in range(4)}, b.out)
        bad.n[0] = 10
        with pytest.raises(PxError):
            b.sim.trace_batch(bad)
        with pytest.raises(PxError):
            b.sim.set_batch_shape(96, 0, 256)
    finally:
        b.gw.close()
