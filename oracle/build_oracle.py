"""Build recipe of the CPU oracle (test infrastructure only).

Compiles px_oracle.c and px_shade.c into oracle/liboracle.so with IEEE single-precision
arithmetic rounded once per operation (no FMA contraction, no fast-math).
The reference (Crunchy000/pixels) is TypeScript + WGSL with no C/C++ sources,
so there is no reference build to place under oracle/_ref.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(verbose: bool = False, out: str | None = None, defines=()) -> str:
    """Builds liboracle.so (or `out`).  `defines` are -D macros; the tests use
    them to plant known mistakes (O_PLANT_*) in a copy built elsewhere."""
    out = out or os.path.join(HERE, "liboracle.so")
    tmp = out + ".tmp"
    cmd = ["gcc", "-O3", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
           "-fexcess-precision=standard", "-pthread"] + ["-D" + d for d in defines] + [
           os.path.join(HERE, "px_oracle.c"), os.path.join(HERE, "px_shade.c"), "-lm", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
