"""Input and output files of the reference's harness.

src/main.cpp reads an input image and a velocity field from PNG files with
libpng's simplified API (includes/utils.hpp:32-104), rescales the velocity
field to [-1, 1] on all four channels (src/main.cpp:170-179), initialises
the temporary velocity buffer to (-1, -1, -1, +1) (src/main.cpp:188-195),
and after every step writes the image as an 8-bit RGBA PNG named <step>.png
(src/main.cpp:59-67, includes/utils.hpp:115-150).  The PNG codec is the
library's own (pfs_png_* in include/pfs.h); besides PNG, the raw float
fixture format here is numpy's .npy, float32 [H][W][4], loaded without
pickles.
"""
from __future__ import annotations

import ctypes
import os
from typing import Tuple

import numpy as np

from . import _lib
from .scenes import Scene, vtmp_init

VP_RANGE = 2.0  # src/main.cpp:26


def png_size(path: str) -> Tuple[int, int]:
    """(width, height) of a PNG file."""
    w, h = ctypes.c_int(), ctypes.c_int()
    _lib.call("pfs_png_size", os.fsencode(path), ctypes.byref(w), ctypes.byref(h))
    return w.value, h.value


def read_png_rgba8(path: str) -> np.ndarray:
    """(H, W, 4) uint8: the bytes libpng's simplified API gives for
    PNG_FORMAT_RGBA (utils.hpp:58-74)."""
    w, h = png_size(path)
    out = np.empty((h, w, 4), dtype=np.uint8)
    _lib.call("pfs_png_read_rgba8", os.fsencode(path), ctypes.c_void_p(out.ctypes.data), out.nbytes)
    return out


def read_png_to_array(path: str) -> np.ndarray:
    """(H, W, 4) float32, each byte / 255.0 (utils.hpp:83-85)."""
    w, h = png_size(path)
    out = np.empty((h, w, 4), dtype=np.float32)
    _lib.call("pfs_png_read", os.fsencode(path), ctypes.c_void_p(out.ctypes.data), out.size)
    return out


def write_png_from_array(path: str, x: np.ndarray) -> None:
    """Write (H, W, 4) floats as an 8-bit RGBA PNG, byte = (png_byte)(x*255.0)
    (utils.hpp:127-131)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    if x.ndim != 3 or x.shape[2] != 4:
        raise ValueError("need an (H, W, 4) array")
    _lib.call("pfs_png_write", os.fsencode(path), ctypes.c_void_p(x.ctypes.data), x.shape[1],
              x.shape[0])


def write_png_rgba8(path: str, x: np.ndarray) -> None:
    x = np.ascontiguousarray(x, dtype=np.uint8)
    if x.ndim != 3 or x.shape[2] != 4:
        raise ValueError("need an (H, W, 4) array")
    _lib.call("pfs_png_write_rgba8", os.fsencode(path), ctypes.c_void_p(x.ctypes.data),
              x.shape[1], x.shape[0])


def bytes_from_floats(x: np.ndarray) -> np.ndarray:
    """The bytes write_png_from_array stores: (png_byte)(x*255.0) as x86-64
    evaluates it (truncation toward zero; outside the int32 range or NaN the
    integer is INT32_MIN; the low byte is kept)."""
    d = np.asarray(x, dtype=np.float32).astype(np.float64) * 255.0
    ok = (d > -2147483649.0) & (d < 2147483648.0)
    t = np.where(ok, np.trunc(np.where(ok, d, 0.0)), -2147483648.0).astype(np.int64)
    return (t & 0xFF).astype(np.uint8)


def rescale_velocity(v: np.ndarray) -> np.ndarray:
    """(v*VP_RANGE) - VP_RANGE/2.0 in double, stored as float, on every
    channel (src/main.cpp:170-179; its transposed index visits every element
    once, so the map is elementwise)."""
    return (np.asarray(v, dtype=np.float32).astype(np.float64) * VP_RANGE
            - VP_RANGE / 2.0).astype(np.float32)


def read_field(path: str) -> np.ndarray:
    """(H, W, 4) float32 from a PNG (byte/255.0) or a .npy raw float file."""
    if path.endswith(".npy"):
        a = np.load(path, allow_pickle=False)
        if a.dtype != np.float32 or a.ndim != 3 or a.shape[2] != 4:
            raise ValueError(f"{path}: need float32 (H, W, 4), got {a.dtype} {a.shape}")
        return np.ascontiguousarray(a)
    return read_png_to_array(path)


def load_scene(input_image: str, velocity_field: str, dt: float, viscosity: float,
               iters: int = 30, steps: int = 1) -> Scene:
    """The four buffers src/main.cpp:126-195 sets up: image from the input
    file, vp = rescaled velocity PNG (a .npy velocity field is taken as the
    rescaled field itself), tmp = (-1, -1, -1, +1), itmp zeroed (the
    reference leaves it uninitialised; it is fully written before use)."""
    image = read_field(input_image)
    if velocity_field.endswith(".npy"):
        vp = read_field(velocity_field)
    else:
        vp = rescale_velocity(read_png_to_array(velocity_field))
    h, w = vp.shape[:2]
    return Scene(vp=vp, tmp=vtmp_init(w, h), image=image, itmp=np.zeros_like(image), dt=dt,
                 visc=viscosity, iters=iters, steps=steps,
                 name=f"{os.path.basename(input_image)}+{os.path.basename(velocity_field)}")
