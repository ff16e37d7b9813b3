"""numpy restatements of the reference's integer back-trace keys, shared by
the CPU pin (tests/test_oracle.py) and the GPU dump test."""
from __future__ import annotations

import numpy as np


def color_keys_numpy(image_shape, vp, dt):
    """Keys of advect_color (src/fluid.cpp:78-113) evaluated with numpy
    float32 scalar-exact operations: nearest velocity cell (vi, vj) =
    ((int)(i*viw), (int)(j*vih)), back-trace x' = i - ((dt/viw)*u)/Wi wrapped
    by fmod(fmod(x', Wi) + Wi, Wi), then (i0, j0, i1, j1)."""
    ih, iw = image_shape
    vh, vw = vp.shape[:2]
    f32 = np.float32
    viw = f32(vw) / f32(iw)
    vih = f32(vh) / f32(ih)
    i = np.arange(iw, dtype=f32)[None, :]
    j = np.arange(ih, dtype=f32)[:, None]
    vi = (i * viw).astype(np.int32)
    vj = (j * vih).astype(np.int32)
    # The reference reads out of bounds if i * viw rounds up to vw; these
    # shapes never do, so the library's clamp does not enter.
    assert vi.max() < vw and vj.max() < vh
    vi, vj = np.broadcast_to(vi, (ih, iw)), np.broadcast_to(vj, (ih, iw))
    u, v = vp[vj, vi, 0], vp[vj, vi, 1]
    xp = i - ((f32(dt) / viw) * u) / f32(iw)
    yp = j - ((f32(dt) / vih) * v) / f32(ih)
    xp = np.fmod(np.fmod(xp, f32(iw)) + f32(iw), f32(iw))
    yp = np.fmod(np.fmod(yp, f32(ih)) + f32(ih), f32(ih))
    i0 = xp.astype(np.int32)
    j0 = yp.astype(np.int32)
    return np.stack([vi, vj, i0, j0, (i0 + 1) % iw, (j0 + 1) % ih], axis=-1)


def bilerp_from_keys(image, keys, image_shape, vp, dt):
    """advect_color's output (src/fluid.cpp:114-124) from the keys, with the
    weights sx = x' - i0, sy = y' - j0 recomputed in float32 as the
    reference does, blended left to right."""
    ih, iw = image_shape
    vh, vw = vp.shape[:2]
    f32 = np.float32
    viw, vih = f32(vw) / f32(iw), f32(vh) / f32(ih)
    i = np.arange(iw, dtype=f32)[None, :]
    j = np.arange(ih, dtype=f32)[:, None]
    vi, vj, i0, j0, i1, j1 = (keys[..., k] for k in range(6))
    xp = i - ((f32(dt) / viw) * vp[vj, vi, 0]) / f32(iw)
    yp = j - ((f32(dt) / vih) * vp[vj, vi, 1]) / f32(ih)
    xp = np.fmod(np.fmod(xp, f32(iw)) + f32(iw), f32(iw))
    yp = np.fmod(np.fmod(yp, f32(ih)) + f32(ih), f32(ih))
    sx = (xp - i0.astype(f32))[..., None]
    sy = (yp - j0.astype(f32))[..., None]
    one = f32(1)
    return ((one - sx) * (one - sy) * image[j0, i0] + sx * (one - sy) * image[j0, i1]
            + (one - sx) * sy * image[j1, i0] + sx * sy * image[j1, i1]).astype(f32)
