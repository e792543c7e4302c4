"""Learned ray-cost binning over the C ABI (px_mlp_*, px_bin_rays,
px_trace_perm_device).

A 16-64-64-1 MLP predicts log2 of a ray's DDA step count; rays are then
sorted, inside windows of consecutive rays, into cost buckets so that a
wavefront holds rays of similar cost.  The plain DDA stays the judge: the
permuted launch traces every ray with the same kernel, so the hits are
bit-identical to the unsorted launch whatever the network predicts.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def _dp(t) -> C.c_void_p:
    return C.c_void_p(int(t.data_ptr()))


class CostModel:
    """Weights + Adam state on the context's device."""

    def __init__(self, gw, seed: int = 1):
        self.gw = gw
        self.lib = gw.lib
        h = C.c_void_p()
        gw.check(self.lib.px_mlp_create(gw.ctx, seed, C.byref(h)))
        self.h = h
        self.n_params = int(self.lib.px_mlp_param_count())

    def close(self) -> None:
        if getattr(self, "h", None):
            self.lib.px_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        p = np.zeros(self.n_params, np.float32)
        self.gw.check(self.lib.px_mlp_get_params(self.h, N.ptr(p, C.c_float)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        self.gw.check(self.lib.px_mlp_set_params(self.h, N.ptr(p, C.c_float)))

    def features(self, kind: int, rays_t, n: int, feat_t) -> None:
        """feat_t: n * 16 bf16 (a torch tensor of >= 32 n bytes)."""
        self.gw.after_torch()
        self.gw.check(self.lib.px_mlp_features(self.gw.ctx, kind, _dp(rays_t), n, _dp(feat_t)))

    def predict(self, feat_t, n: int, pred_t) -> None:
        self.gw.after_torch()
        self.gw.check(self.lib.px_mlp_predict(self.h, _dp(feat_t), n, _dp(pred_t)))

    def grad(self, feat_t, target_t, n: int):
        self.gw.after_torch()
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float(0.0)
        self.gw.check(self.lib.px_mlp_grad(self.h, _dp(feat_t), _dp(target_t), n, N.ptr(g, C.c_float),
                                           C.byref(loss)))
        return float(loss.value), g

    def sort_rays(self, kind: int, rays_t, n: int, perm_t, buckets: int = 32, lo: float = 0.0,
                  hi: float = 9.0, faces: bool = False, window_log2: int = 0) -> None:
        """Features, prediction and a stable sort of all n rays by cost
        bucket (three launches), or by (direction face, cost bucket) with
        faces=True; with window_log2, each window of 2**window_log2
        consecutive rays is sorted on its own.  Without faces or windows,
        perm_t gets the order bin_rays(window=n) gives for
        predict(features(rays))."""
        self.gw.after_torch()
        flags = (N.SORT_FACES if faces else 0) | (N.SORT_WINDOW(window_log2) if window_log2 else 0)
        self.gw.check(self.lib.px_mlp_sort_rays(self.h, kind, _dp(rays_t), n, buckets, lo, hi, flags, _dp(perm_t)))

    def trace_sorted(self, kind: int, rays_t, n: int, perm_t, out_t, buckets: int = 32, lo: float = 0.0,
                     hi: float = 9.0, faces: bool = False, window_log2: int = 0) -> None:
        """sort_rays, then the permuted launch, in one call (one submission
        of four kernels); out_t gets trace_device's hits bit for bit."""
        self.gw.after_torch()
        flags = (N.SORT_FACES if faces else 0) | (N.SORT_WINDOW(window_log2) if window_log2 else 0)
        self.gw.check(self.lib.px_mlp_trace_sorted(self.h, kind, _dp(rays_t), n, buckets, lo, hi, flags,
                                                   _dp(perm_t), _dp(out_t)))

    def train_step(self, feat_t, target_t, n: int, lr: float = 1e-3, want_loss: bool = False):
        """One Adam step, asynchronous on the library's stream unless
        want_loss: keep feat_t / target_t alive (or call gw.sync()) until it
        has run, since torch may reuse their memory as soon as they are
        freed."""
        self.gw.after_torch()
        loss = C.c_float(0.0)
        self.gw.check(self.lib.px_mlp_train_step(self.h, _dp(feat_t), _dp(target_t), n, lr,
                                                 C.byref(loss) if want_loss else None))
        return float(loss.value) if want_loss else None


def bin_rays(gw, pred_t, n: int, perm_t, window: int = 256, buckets: int = 32, lo: float = 0.0, hi: float = 9.0):
    gw.after_torch()
    gw.check(gw.lib.px_bin_rays(gw.ctx, _dp(pred_t), n, window, buckets, lo, hi, _dp(perm_t)))


def trace_perm(gw, kind: int, rays_t, perm_t, n: int, out_t) -> None:
    gw.after_torch()
    gw.check(gw.lib.px_trace_perm_device(gw.ctx, kind, _dp(rays_t), _dp(perm_t), n, _dp(out_t)))


def log2_steps(steps: np.ndarray) -> np.ndarray:
    """Training target: log2 of the step count (0 for rays with no step)."""
    return np.log2(np.maximum(np.asarray(steps, np.float64), 1.0)).astype(np.float32)


def wave_cost(steps_in_order: np.ndarray, wave: int = 64) -> int:
    """Sum over wavefronts of the longest ray's step count: the issue cost of
    a launch whose lanes wait for the slowest lane."""
    s = np.asarray(steps_in_order, np.int64)
    pad = (-s.size) % wave
    if pad:
        s = np.concatenate([s, np.zeros(pad, np.int64)])
    return int(s.reshape(-1, wave).max(1).sum())
