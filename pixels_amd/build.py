"""Build the native libraries of pixels_amd in-tree.

  libpixels_hip.so  HIP kernels + C ABI (include/pixels.h), gfx950 only
  libpxscene.so     seeded scene builder (host C++)

Run as ``python -m pixels_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ROOT = os.path.dirname(HERE)

HIP_SOURCES = ["px_api.hip", "px_trace.hip", "px_edit.hip", "px_mlp.hip", "px_particles.hip", "px_sort.hip", "px_scene.hip", "px_render.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    # packed FP32 instructions stay off (see the build notes in DESIGN.md)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))


def build_hip(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(HERE, "libpixels_hip.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_scene(verbose: bool = False) -> str:
    out = os.path.join(HERE, "libpxscene.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-pthread",
          os.path.join(CSRC, "scene.cpp"), "-o", tmp])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


NODE_INCLUDE = "/usr/include/node"


def build_napi(verbose: bool = False) -> str | None:
    """N-API addon for the Node host (needs node_api.h); links libpixels_hip.so."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        if verbose:
            print("node_api.h not found: N-API addon skipped")
        return None
    out = os.path.join(HERE, "node", "pixels_napi.node")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + NODE_INCLUDE, "-DNAPI_VERSION=8",
          os.path.join(HERE, "node", "px_napi.cc"), "-o", tmp, "-L" + HERE, "-lpixels_hip",
          "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_all(verbose: bool = False) -> None:
    with ThreadPoolExecutor(max_workers=2) as ex:
        fs = [ex.submit(build_hip, verbose), ex.submit(build_scene, verbose)]
        for f in fs:
            f.result()
    build_napi(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
