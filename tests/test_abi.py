"""The C-ABI library loads and exports every entry point include/pixels.h
declares (no compute calls: this runs without a GPU)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(px_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_the_boundary():
    names = declared_symbols()
    for n in ("px_create", "px_destroy", "px_trace", "px_trace_device", "px_trace_frame_device", "px_probe",
              "px_read_ctrl", "px_push_impacts", "px_island_pass", "px_island_labels", "px_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import ctypes

    from pixels_amd import _native
    lib = _native.load()
    for n in sorted(declared_symbols()):
        assert hasattr(lib, n), n
        assert isinstance(getattr(lib, n), ctypes._CFuncPtr)
    assert set(_native.SIGNATURES) == declared_symbols()


def test_default_limits_are_the_reference_limits():
    from pixels_amd import default_limits
    lim = default_limits()
    # src/gpu/layout.ts:11-33
    assert (lim.max_dirty, lim.max_alloc, lim.max_writes, lim.max_impacts, lim.max_jobs) == (65535, 65535, 131072, 64,
                                                                                               65535)
    assert (lim.max_flood, lim.max_seeds, lim.hash_size, lim.component_budget) == (131072, 262144, 1 << 20, 400000)
    assert (lim.search_iters_per_frame, lim.flood_iters_per_frame) == (16, 12)


def test_create_without_gpu_fails_cleanly():
    from tests.conftest import has_gpu
    if has_gpu():
        pytest.skip("a GPU is present")
    import numpy as np

    from pixels_amd import GpuWorld, PxError, Scene
    s = Scene.from_dense(np.zeros((8, 8, 8), np.uint16))
    with pytest.raises(PxError):
        GpuWorld(s)


def test_ctrl_names_match_header_enum():
    from pixels_amd import CTRL_NAMES
    text = open(os.path.join(ROOT, "include", "pixels.h")).read()
    body = text[text.index("enum px_ctrl_word"):]
    body = body[:body.index("};")]
    names = re.findall(r"PX_([CS]_[A-Z_]+)", body)
    assert names == CTRL_NAMES
