"""Autotuner for the traversal launch shape.

Runs the frame (or a ray batch) under candidate shapes of the persistent
kernel - workgroups launched, threads per workgroup, rays claimed per
wavefront atomic (chunk), idle lanes that trigger a refill - times each with
the library's HIP-event profiling, and keeps the fastest.  Outputs do not
depend on the shape (every shape is bit-identical), so tuning only moves time.
"""
from __future__ import annotations

import itertools

from . import _native as N

# CDNA4: 256 CUs, 4 SIMDs each; the persistent kernel holds 5 waves per SIMD
DEFAULT_SPACE = {
    "block": (64, 128, 256),
    "waves_per_cu": (8, 16, 20, 32),
    "chunk": (64, 256, 1024),
    "refill_min": (8, 16, 32),
}


def _candidates(space, cus=256):
    for block, wpc, chunk, refill in itertools.product(space["block"], space["waves_per_cu"], space["chunk"],
                                                        space["refill_min"]):
        waves = cus * wpc
        blocks = max(1, waves * 64 // block)
        yield dict(variant=1, blocks=blocks, block=block, chunk=chunk, refill_min=refill)


def tune_frame(gw, sim, run_once, space=None, repeats: int = 2, coarse: bool = True, log=None):
    """run_once(): launches one frame on the context's stream.  Returns the
    best shape (dict) and a list of (shape, ms) results; the best shape is
    left set on the context."""
    space = dict(DEFAULT_SPACE, **(space or {}))
    results = []

    def measure(shape):
        gw.set_tuning(**shape)
        run_once()  # warm
        gw.sync()
        gw.profile_begin()
        for _ in range(repeats):
            run_once()
        prof = gw.profile_end()
        ms = sum(v[0] for k, v in prof.items() if k in (N.PROF_PRIMARY, N.PROF_SHADOW, N.PROF_SHADOW_GEN,
                                                         N.PROF_TRACE)) / repeats
        results.append((dict(shape), ms))
        if log:
            log("tune %s: %.3f ms" % (shape, ms))
        return ms

    best = dict(variant=0, blocks=0, block=0, chunk=0, refill_min=0)
    best_ms = measure(best)
    if coarse:
        # one axis at a time around a sensible start
        cur = dict(variant=1, blocks=256 * 20 * 64 // 256, block=256, chunk=256, refill_min=16)
        cur_ms = measure(cur)
        for axis in ("waves_per_cu", "block", "chunk", "refill_min"):
            for val in space[axis]:
                cand = dict(cur)
                if axis == "waves_per_cu":
                    cand["blocks"] = max(1, 256 * val * 64 // cand["block"])
                elif axis == "block":
                    waves = cand["blocks"] * cand["block"] // 64
                    cand["block"] = val
                    cand["blocks"] = max(1, waves * 64 // val)
                else:
                    cand[axis] = val
                if cand == cur:
                    continue
                ms = measure(cand)
                if ms < cur_ms:
                    cur, cur_ms = cand, ms
        if cur_ms < best_ms:
            best, best_ms = cur, cur_ms
    else:
        for cand in _candidates(space):
            ms = measure(cand)
            if ms < best_ms:
                best, best_ms = cand, ms
    gw.set_tuning(**best)
    return best, best_ms, results
