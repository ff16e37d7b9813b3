"""MI355X (gfx950) implementation of the reference's grid fluid step.

Drop-in for the functions of the reference's includes/fluid.hpp (see
``fluid``), backed by hand-written HIP kernels in libpfs.so (``csrc/``),
with results identical to the reference's CPU solver src/fluid.cpp.
"""
from __future__ import annotations

__all__ = ["fluid", "scenes", "build"]


def build(verbose: bool = False) -> str:
    """Compile libpfs.so in-tree for gfx950; returns its path."""
    import os
    import subprocess

    here = os.path.dirname(os.path.abspath(__file__))
    cmd = ["make", "-C", os.path.join(here, "csrc"), "-j8"]
    out = subprocess.run(cmd, capture_output=not verbose, text=True)
    if out.returncode != 0:
        raise RuntimeError("libpfs build failed:\n" + (out.stdout or "") + (out.stderr or ""))
    return os.path.join(here, "libpfs.so")
