"""Deterministic synthetic inputs for the fluid step.

The reference loads its velocity field and dye image from PNG files
(``src/main.cpp:129-140``) and then initialises the buffers it hands to the
solver (``src/main.cpp:170-195``):

* ``vp`` = PNG/255 rescaled ``2v - 1`` on all four channels, so a velocity PNG
  with blue 0 and alpha 255 gives ``(u, v, -1, +1)`` per cell;
* ``vtmp`` = ``(-1, -1, -1, +1)`` per cell (``src/main.cpp:188-195``);
* ``itmp`` is left uninitialised (every channel is written before it is read).

The generators here produce those same buffers from seeded synthetic fields
instead of files, so the CPU oracle and the GPU receive byte-identical
arrays.  Randomness comes from SplitMix64 (Steele, Lea & Flood, 2014) and
normal draws from the Box-Muller transform; both are evaluated in float64
numpy and rounded once to float32.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n outputs of SplitMix64 started at state = seed + stream * 2^32 (uint64)."""
    base = np.uint64((int(seed) + (int(stream) << 32)) & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        z = base + _GOLDEN * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n float64 draws in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n N(0,1) draws by Box-Muller over pairs of uniforms."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, stream)
    r = np.sqrt(-2.0 * np.log1p(-u[0::2]))
    t = 2.0 * math.pi * u[1::2]
    return np.concatenate([r * np.cos(t), r * np.sin(t)])[:n]


def _grid(w: int, h: int):
    x = (np.arange(w, dtype=np.float64) + 0.5) / w
    y = (np.arange(h, dtype=np.float64) + 0.5) / h
    return np.meshgrid(x, y)  # (h, w) each


def taylor_green(w: int, h: int, n: int, seed: int, amp=(0.1, 1.0)):
    """Sum of n periodic Taylor-Green cells with random wavenumber 1..4,
    phases in [0, 2pi) and amplitudes in ``amp``; normalised to max |u|,|v| = 1."""
    r = uniform(seed, 5 * n, stream=1).reshape(n, 5)
    X, Y = _grid(w, h)
    u = np.zeros((h, w))
    v = np.zeros((h, w))
    for k in range(n):
        kw = 1 + int(r[k, 0] * 4)
        phx, phy = 2 * math.pi * r[k, 1], 2 * math.pi * r[k, 2]
        a = amp[0] + (amp[1] - amp[0]) * r[k, 3]
        sx, sy = 2 * math.pi * kw * X + phx, 2 * math.pi * kw * Y + phy
        u += a * np.sin(sx) * np.cos(sy)
        v += -a * np.cos(sx) * np.sin(sy)
    s = max(np.abs(u).max(), np.abs(v).max(), 1e-30)
    return u / s, v / s


def point_vortices(w: int, h: int, n: int, seed: int, noise: float = 0.0):
    """n Gaussian-core vortices (periodic minimum image) with random centre,
    core radius in [0.02, 0.1] and signed strength, normalised to 0.95, plus
    N(0, noise) per component, clipped to [-1, 1]."""
    r = uniform(seed, 4 * n, stream=2).reshape(n, 4)
    X, Y = _grid(w, h)
    u = np.zeros((h, w))
    v = np.zeros((h, w))
    for k in range(n):
        cx, cy = r[k, 0], r[k, 1]
        sig = 0.02 + 0.08 * r[k, 2]
        g = 2.0 * r[k, 3] - 1.0
        dx = (X - cx + 0.5) % 1.0 - 0.5
        dy = (Y - cy + 0.5) % 1.0 - 0.5
        e = np.exp(-(dx * dx + dy * dy) / (2 * sig * sig)) / sig
        u += -g * dy * e
        v += g * dx * e
    s = max(np.abs(u).max(), np.abs(v).max(), 1e-30)
    u, v = 0.95 * u / s, 0.95 * v / s
    if noise > 0:
        z = normal(seed, 2 * w * h, stream=3).reshape(2, h, w)
        u = u + noise * z[0]
        v = v + noise * z[1]
    return np.clip(u, -1.0, 1.0), np.clip(v, -1.0, 1.0)


def dye_disk(w: int, h: int, radius: float = 0.3) -> np.ndarray:
    """RGBA dye: R = 1 inside a centred disk of the given radius (unit square), G and
    B linear ramps in x and y, A = 1."""
    X, Y = _grid(w, h)
    img = np.empty((h, w, 4))
    img[..., 0] = ((X - 0.5) ** 2 + (Y - 0.5) ** 2 <= radius * radius)
    img[..., 1] = X
    img[..., 2] = Y
    img[..., 3] = 1.0
    return img.astype(np.float32)


def dye_blobs(w: int, h: int, n: int, seed: int, sigma_px=(5.0, 40.0), thresh: float = 0.5):
    """RGBA dye: R = [sum of n unit Gaussian blobs (sigma in pixels) >= thresh],
    G = the unthresholded sum clipped to [0, 1], B = 1 - R, A = 1."""
    r = uniform(seed, 3 * n, stream=4).reshape(n, 3)
    xs = np.arange(w, dtype=np.float64) + 0.5
    ys = np.arange(h, dtype=np.float64) + 0.5
    acc = np.zeros((h, w))
    for k in range(n):
        cx, cy = r[k, 0] * w, r[k, 1] * h
        s = sigma_px[0] + (sigma_px[1] - sigma_px[0]) * r[k, 2]
        ex = np.exp(-(((xs - cx + w / 2) % w - w / 2) ** 2) / (2 * s * s))
        ey = np.exp(-(((ys - cy + h / 2) % h - h / 2) ** 2) / (2 * s * s))
        acc += ey[:, None] * ex[None, :]
    mask = (acc >= thresh).astype(np.float64)
    img = np.empty((h, w, 4))
    img[..., 0] = mask
    img[..., 1] = np.clip(acc, 0.0, 1.0)
    img[..., 2] = 1.0 - mask
    img[..., 3] = 1.0
    return img.astype(np.float32)


def vp_from_uv(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """(H, W, 4) float32 velocity/pressure field (u, v, -1, +1), the state
    src/main.cpp:170-179 produces from a velocity PNG with B = 0, A = 255."""
    h, w = u.shape
    vp = np.empty((h, w, 4), dtype=np.float32)
    vp[..., 0] = u.astype(np.float32)
    vp[..., 1] = v.astype(np.float32)
    vp[..., 2] = -1.0
    vp[..., 3] = 1.0
    return vp


def vtmp_init(w: int, h: int) -> np.ndarray:
    """(-1, -1, -1, +1) per cell, src/main.cpp:188-195."""
    t = np.full((h, w, 4), -1.0, dtype=np.float32)
    t[..., 3] = 1.0
    return t


@dataclasses.dataclass
class Scene:
    """One simulation job: the four buffers of src/main.cpp plus step parameters."""
    vp: np.ndarray
    tmp: np.ndarray
    image: np.ndarray
    itmp: np.ndarray
    dt: float
    visc: float
    iters: int
    steps: int
    name: str = ""

    @property
    def cells(self) -> int:
        return self.vp.shape[0] * self.vp.shape[1]

    @property
    def pixels(self) -> int:
        return self.image.shape[0] * self.image.shape[1]

    def copy(self) -> "Scene":
        return dataclasses.replace(self, vp=self.vp.copy(), tmp=self.tmp.copy(),
                                   image=self.image.copy(), itmp=self.itmp.copy())


def make_scene(fw: int, fh: int, iw: int, ih: int, *, kind: str = "tg", seed: int = 0,
               dt: float = 1.0 / 60.0, visc: float = 0.1, iters: int = 4, steps: int = 1,
               vel_scale: float = 1.0, name: str = "") -> Scene:
    """Build a scene.  ``kind`` "tg" = 4 Taylor-Green cells + disk dye (the
    16K-cell CPU-runnable workload); "vortex" = 16 point vortices + N(0, 0.05)
    noise + 64 thresholded Gaussian blobs (the 1M-cell workload).  ``vel_scale``
    multiplies u, v (used by tests to force multi-cell back-traces)."""
    if kind == "tg":
        u, v = taylor_green(fw, fh, 4, seed)
        image = dye_disk(iw, ih, 0.3)
    elif kind == "vortex":
        u, v = point_vortices(fw, fh, 16, seed, noise=0.05)
        image = dye_blobs(iw, ih, 64, seed)
    else:
        raise ValueError(kind)
    vp = vp_from_uv(u * vel_scale, v * vel_scale)
    return Scene(vp=vp, tmp=vtmp_init(fw, fh), image=image,
                 itmp=np.zeros_like(image), dt=dt, visc=visc, iters=iters, steps=steps,
                 name=name or f"{kind}{fw}x{fh}")


def random_scene(fw: int, fh: int, iw: int, ih: int, seed: int, *, dt: float = 1.0,
                 visc: float = 0.37, iters: int = 4, vel_scale: Optional[float] = None,
                 channels: int = 4) -> Scene:
    """Uniform-random state (every channel random, incl. pressure and the dye),
    with velocities large enough that back-traces cross several cells.  For
    parity tests: exercises every gather, wrap and stale-channel path."""
    u = uniform(seed, 4 * fw * fh + 2 * fw * fh, stream=5)
    vs = vel_scale if vel_scale is not None else 3.0 * max(fw, fh)
    vp = np.empty((fh, fw, 4), dtype=np.float32)
    vp[..., 0] = ((2 * u[0::6][: fw * fh] - 1) * vs).reshape(fh, fw)
    vp[..., 1] = ((2 * u[1::6][: fw * fh] - 1) * vs).reshape(fh, fw)
    vp[..., 2] = (2 * u[2::6][: fw * fh] - 1).reshape(fh, fw)
    vp[..., 3] = (2 * u[3::6][: fw * fh] - 1).reshape(fh, fw)
    tmp = np.empty_like(vp)
    tmp[..., 0] = (2 * u[4::6][: fw * fh] - 1).reshape(fh, fw)
    tmp[..., 1] = (2 * u[5::6][: fw * fh] - 1).reshape(fh, fw)
    tmp[..., 2] = (u[4::6][: fw * fh] - 0.5).reshape(fh, fw)
    tmp[..., 3] = (u[5::6][: fw * fh] - 0.5).reshape(fh, fw)
    im = uniform(seed, ih * iw * channels, stream=6).astype(np.float32).reshape(ih, iw, channels)
    return Scene(vp=vp, tmp=tmp, image=im, itmp=np.zeros_like(im), dt=dt, visc=visc,
                 iters=iters, steps=1, name=f"rand{fw}x{fh}/{iw}x{ih}s{seed}")


# --- workloads named by their shapes --------------------------------------------
def plain_16k(steps: int = 200) -> Scene:
    """128x128 field and dye, 4 Taylor-Green cells (seed 0), disk r = 0.3,
    K = 4, dt = 1/60, nu = 0.1."""
    return make_scene(128, 128, 128, 128, kind="tg", seed=0, dt=1.0 / 60.0, visc=0.1,
                      iters=4, steps=steps, name="plain-128x128-K4")


def plain_1m(steps: int = 500) -> Scene:
    """1024x1024 field and dye, 16 vortices + N(0,0.05) noise, 64 thresholded
    blobs (seed 0), K = 8, dt = 1/120, nu = 0.1."""
    return make_scene(1024, 1024, 1024, 1024, kind="vortex", seed=0, dt=1.0 / 120.0,
                      visc=0.1, iters=8, steps=steps, name="plain-1024x1024-K8")


def precision_4m(steps: int = 200) -> Scene:
    """2048x2048 field and dye with the 16K workload's generator, K <= 16."""
    return make_scene(2048, 2048, 2048, 2048, kind="tg", seed=0, dt=1.0 / 60.0, visc=0.1,
                      iters=16, steps=steps, name="precision-2048x2048-K16")


def batched_specs(n: int = 256, seed: int = 0) -> List[dict]:
    """Scene shapes of the batched workload: dye pixels log-uniform in
    [4096, 262144] (square side = round(sqrt)), field side uniform in
    {64, 128, 256, 512}; scene s is generated from seed s."""
    r = uniform(seed, 2 * n, stream=7).reshape(n, 2)
    out = []
    for s in range(n):
        m = math.exp(math.log(4096) + (math.log(262144) - math.log(4096)) * r[s, 0])
        side = max(1, int(round(math.sqrt(m))))
        f = (64, 128, 256, 512)[min(3, int(r[s, 1] * 4))]
        out.append(dict(scene=s, fw=f, fh=f, iw=side, ih=side, seed=s))
    return out


def served_specs(n: int, rate: float, seed: int = 7) -> List[dict]:
    """Job stream of the served workload: Poisson arrivals at ``rate`` jobs/s,
    dye pixels log-uniform in [8192, 2^20], field side = round(sqrt(pixels)),
    steps uniform in [50, 500]."""
    r = uniform(seed, 3 * n, stream=8).reshape(n, 3)
    t = 0.0
    out = []
    for k in range(n):
        t += -math.log1p(-r[k, 0]) / rate
        m = math.exp(math.log(8192) + (math.log(2 ** 20) - math.log(8192)) * r[k, 1])
        side = max(1, int(round(math.sqrt(m))))
        steps = 50 + int(r[k, 2] * 451)
        out.append(dict(job=k, arrival=t, fw=side, fh=side, iw=side, ih=side, steps=steps,
                        seed=k))
    return out


# ---- particle sets (the particle layer; no reference counterpart) -----------
def particles_disk(n: int, w: int, h: int, seed: int = 0, radius: float = 0.3):
    """n positions uniform in the disk of radius `radius` (domain units)
    centred in the domain, in field-cell units: (x, y) float32 arrays."""
    r = uniform(seed, 2 * n, stream=6).reshape(n, 2)
    rad = radius * np.sqrt(r[:, 0])
    ang = 2.0 * np.pi * r[:, 1]
    x = (0.5 + rad * np.cos(ang)) * w
    y = (0.5 + rad * np.sin(ang)) * h
    return x.astype(np.float32), y.astype(np.float32)


def particles_in_mask(mask: np.ndarray, n: int, seed: int = 0):
    """n positions uniform over the set pixels of an (H, W) mask: a set pixel
    drawn uniformly, then a uniform point inside it (field-cell units)."""
    idx = np.flatnonzero(mask.reshape(-1) > 0)
    if idx.size == 0:
        raise ValueError("empty mask")
    h, w = mask.shape
    r = uniform(seed, 3 * n, stream=7).reshape(n, 3)
    pick = idx[np.minimum((r[:, 0] * idx.size).astype(np.int64), idx.size - 1)]
    x = ((pick % w) + r[:, 1]).astype(np.float32)
    y = ((pick // w) + r[:, 2]).astype(np.float32)
    # float32 rounding may reach the far edge: keep positions inside [0, W)
    x = np.minimum(x, np.nextafter(np.float32(w), np.float32(0)))
    y = np.minimum(y, np.nextafter(np.float32(h), np.float32(0)))
    return x, y


def particles_16k():
    """16,384 particles in the disk of radius 0.3 on the 128x128 Taylor-Green
    field of the 16K workload; radius 0.02 of the domain (2.56 cells)."""
    s = plain_16k()
    x, y = particles_disk(16384, 128, 128, seed=0)
    return s, x, y, 0.02 * 128


def particles_1m():
    """1,048,576 particles over the set pixels of the 1M workload's dye blobs,
    on its 1024x1024 vortex field; radius 1.5 cells."""
    s = plain_1m()
    x, y = particles_in_mask(s.image[..., 0], 1 << 20, seed=0)
    return s, x, y, 1.5
