"""The reference's PNG reader and writer, run for real — TEST INFRASTRUCTURE ONLY.

includes/utils.hpp:32-150 reads every input with libpng's simplified API
(png_image_begin_read_from_file, format PNG_FORMAT_RGBA,
png_image_finish_read with no background and the default row stride) and
writes frames with png_image_write_to_file.  libpng is a third-party
dependency of the reference (vcpkg.json) and is not part of its tree; this
module calls the system's libpng16 shared library through its public
simplified API, exactly as utils.hpp does, so that the library's own PNG
codec (probabilistic_fluid_simulation_amd/csrc/pfs_png.cpp) can be checked
against it byte for byte.  It is the oracle for that codec, never used by
it.  ``available()`` is False where libpng16 is not installed.

The png_image layout and the constants below are libpng 1.6's public API:
PNG_IMAGE_VERSION 1, PNG_FORMAT_RGBA = COLOR | ALPHA = 3.
"""
from __future__ import annotations

import ctypes
import ctypes.util
from typing import Optional

import numpy as np

PNG_IMAGE_VERSION = 1
PNG_FORMAT_RGBA = 3


class _PngImage(ctypes.Structure):
    _fields_ = [("opaque", ctypes.c_void_p), ("version", ctypes.c_uint32),
                ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("format", ctypes.c_uint32), ("flags", ctypes.c_uint32),
                ("colormap_entries", ctypes.c_uint32), ("warning_or_error", ctypes.c_uint32),
                ("message", ctypes.c_char * 64)]


_lib = None
_tried = False


def _load() -> Optional[ctypes.CDLL]:
    global _lib, _tried
    if _tried:
        return _lib
    _tried = True
    for name in ("libpng16.so.16", ctypes.util.find_library("png16") or ""):
        if not name:
            continue
        try:
            lib = ctypes.CDLL(name)
        except OSError:
            continue
        P = ctypes.POINTER(_PngImage)
        lib.png_get_libpng_ver.restype = ctypes.c_char_p
        lib.png_get_libpng_ver.argtypes = [ctypes.c_void_p]
        lib.png_image_begin_read_from_file.restype = ctypes.c_int
        lib.png_image_begin_read_from_file.argtypes = [P, ctypes.c_char_p]
        lib.png_image_finish_read.restype = ctypes.c_int
        lib.png_image_finish_read.argtypes = [P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                              ctypes.c_void_p]
        lib.png_image_free.restype = None
        lib.png_image_free.argtypes = [P]
        lib.png_image_write_to_file.restype = ctypes.c_int
        lib.png_image_write_to_file.argtypes = [P, ctypes.c_char_p, ctypes.c_int, ctypes.c_void_p,
                                                ctypes.c_int32, ctypes.c_void_p]
        _lib = lib
        break
    return _lib


def available() -> bool:
    return _load() is not None


def version() -> str:
    lib = _load()
    return lib.png_get_libpng_ver(None).decode() if lib else ""


def read_rgba8(path: str) -> np.ndarray:
    """(H, W, 4) uint8 as utils.hpp:50-74 obtains it; raises ValueError with
    libpng's message if libpng rejects the file."""
    lib = _load()
    if lib is None:
        raise RuntimeError("libpng16 is not installed")
    im = _PngImage()
    im.version = PNG_IMAGE_VERSION
    if not lib.png_image_begin_read_from_file(ctypes.byref(im), path.encode()):
        raise ValueError(im.message.decode(errors="replace"))
    im.format = PNG_FORMAT_RGBA
    out = np.zeros((im.height, im.width, 4), dtype=np.uint8)
    ok = lib.png_image_finish_read(ctypes.byref(im), None, ctypes.c_void_p(out.ctypes.data), 0,
                                   None)
    msg = im.message.decode(errors="replace")
    lib.png_image_free(ctypes.byref(im))
    if not ok:
        raise ValueError(msg)
    return out


def write_rgba8(path: str, rgba: np.ndarray) -> None:
    """png_image_write_to_file of an 8-bit RGBA buffer (utils.hpp:134)."""
    lib = _load()
    if lib is None:
        raise RuntimeError("libpng16 is not installed")
    rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
    im = _PngImage()
    im.version = PNG_IMAGE_VERSION
    im.width, im.height = rgba.shape[1], rgba.shape[0]
    im.format = PNG_FORMAT_RGBA
    if not lib.png_image_write_to_file(ctypes.byref(im), path.encode(), 0,
                                       ctypes.c_void_p(rgba.ctypes.data), 0, None):
        raise ValueError(im.message.decode(errors="replace"))
