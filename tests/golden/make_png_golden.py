"""Writes tests/golden/png_libpng.npz: the RGBA8 bytes that libpng's
simplified read API (the reference's PNG reader, includes/utils.hpp:32-104)
produces for every decoding case of tests/png_util.cases(0), so that the
library's PNG codec stays pinned where libpng is not installed.  For
interlaced 16-bit cases libpng decodes the same pixels written
non-interlaced (png_util.spec_equivalent: libpng 1.6's simplified reader
mis-assembles interlaced 16-bit images).

    python tests/golden/make_png_golden.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import png_util  # noqa: E402
from oracle import libpng_ref  # noqa: E402


def main():
    if not libpng_ref.available():
        raise SystemExit("libpng16 is not installed")
    out = {}
    with tempfile.TemporaryDirectory() as d:
        for name, kw in png_util.cases(0):
            path = png_util.write(os.path.join(d, name + ".png"), **png_util.spec_equivalent(kw))
            out[name] = libpng_ref.read_rgba8(path)
    out["_libpng_version"] = np.frombuffer(libpng_ref.version().encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "png_libpng.npz"), **out)
    print(f"{len(out) - 1} cases, libpng {libpng_ref.version()}")


if __name__ == "__main__":
    main()
