"""px_trace_frame_device_noalloc (the frame with its shadow-ray queue in the
library's fixed scratch array, no allocation) against px_trace_frame_device
on the same inputs, through the C ABI: two calls in a row, the second on a
smaller ray set (so stale queue entries of the first call would show), in
both output layouts, bit for bit, cast counts included; then the
capacity check and the empty frame."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(gw, fn, rays, n, slots):
    import torch
    dr = torch.from_numpy(rays[:n].view(np.uint8).copy()).cuda()
    dp = torch.full((n * 16,), 0xAB, dtype=torch.uint8, device="cuda")
    ds = torch.full((n * slots * 16,), 0xCD, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    nc = C.c_uint32(0)
    gw.check(fn(gw.ctx, C.c_void_p(dr.data_ptr()), n, C.c_void_p(dp.data_ptr()), C.c_void_p(ds.data_ptr()),
                C.byref(nc)))
    gw.sync()
    return dp.cpu().numpy(), ds.cpu().numpy(), int(nc.value)


@pytest.mark.parametrize("planar", [False, True])
def test_noalloc_frame_equals_frame(small_scene, planar):
    from pixels_amd import GpuWorld, Simulation
    s = small_scene
    gw = GpuWorld(s)
    try:
        sim = Simulation(gw)
        sim.set_frame_layout(planar)
        lib = gw.lib
        for (W, H, seed, cut) in ((640, 360, 4, 0), (200, 120, 6, 37)):
            rays = s.camera_rays(W, H, 90.0, seed=seed)
            u, lights = s.frame_inputs(W, H, seed=seed)
            gw.set_uniforms(u)
            gw.set_lights(lights)
            slots = sim.frame_slots()
            n = W * H - cut
            print("case: noalloc frame n=%d planar=%s" % (n, planar), flush=True)
            a = _run(gw, lib.px_trace_frame_device_noalloc, rays, n, slots)
            b = _run(gw, lib.px_trace_frame_device, rays, n, slots)
            assert np.array_equal(a[0], b[0]), "primary hits differ"
            assert np.array_equal(a[1], b[1]), "shadow hits differ"
            assert a[2] == b[2] and a[2] > 0
    finally:
        gw.close()


def test_noalloc_limits(small_scene):
    from pixels_amd import GpuWorld, PxError, Simulation
    s = small_scene
    gw = GpuWorld(s)
    try:
        sim = Simulation(gw)
        u, lights = s.frame_inputs(64, 64, seed=1)
        gw.set_uniforms(u)
        gw.set_lights(lights)
        slots = sim.frame_slots()
        # 64 * ceil(n / 64) * slots words beyond the 20,971,520-word array: refused on the host
        n_big = ((20 << 20) // (64 * slots) + 1) * 64
        with pytest.raises(PxError):
            gw.check(gw.lib.px_trace_frame_device_noalloc(gw.ctx, C.c_void_p(16), n_big, C.c_void_p(16),
                                                          C.c_void_p(16), None))
        assert "fixed scratch" in gw.lib.px_last_error(gw.ctx).decode()
        nc = C.c_uint32(7)
        gw.check(gw.lib.px_trace_frame_device_noalloc(gw.ctx, None, 0, None, None, C.byref(nc)))
        assert nc.value == 0
    finally:
        gw.close()
