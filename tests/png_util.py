"""A minimal PNG writer for test inputs (any colour type and bit depth, chosen
row filters, Adam7 interlace, extra chunks), written from the PNG
specification (ISO/IEC 15948).  Test infrastructure only."""
from __future__ import annotations

import struct
import zlib
from typing import Iterable, List, Sequence, Tuple

import numpy as np

CHANNELS = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}
ADAM7 = [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
         (0, 1, 1, 2)]


def chunk(t: bytes, d: bytes) -> bytes:
    return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)


def pack_row(samples: np.ndarray, depth: int) -> bytes:
    """samples: (W, C) ints."""
    flat = samples.reshape(-1).astype(np.int64)
    if depth == 16:
        return flat.astype(">u2").tobytes()
    if depth == 8:
        return flat.astype(np.uint8).tobytes()
    per = 8 // depth
    out = bytearray((len(flat) * depth + 7) // 8)
    for k, v in enumerate(flat):
        out[k // per] |= (int(v) & ((1 << depth) - 1)) << (8 - depth * (k % per + 1))
    return bytes(out)


def _paeth(a: int, b: int, c: int) -> int:
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if pa <= pb and pa <= pc else (b if pb <= pc else c)


def filter_row(raw: bytes, prev: bytes, ft: int, bpp: int) -> bytes:
    out = bytearray(len(raw))
    for i, x in enumerate(raw):
        a = raw[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        pred = (0, a, b, (a + b) >> 1, _paeth(a, b, c))[ft]
        out[i] = (x - pred) & 0xFF
    return bytes([ft]) + bytes(out)


def encode(samples: np.ndarray, depth: int, ctype: int, interlace: bool = False,
           filters: Sequence[int] = (0,), extra: Iterable[Tuple[bytes, bytes]] = (),
           idat_parts: int = 1) -> bytes:
    """samples: (H, W, C) ints (palette images: indices)."""
    h, w, c = samples.shape
    assert c == CHANNELS[ctype]
    bpp = max(1, c * depth // 8)
    passes = ADAM7 if interlace else [(0, 0, 1, 1)]
    raw = bytearray()
    k = 0
    for x0, y0, dx, dy in passes:
        sub = samples[y0::dy, x0::dx]
        if sub.size == 0:
            continue
        rows = [pack_row(r, depth) for r in sub]
        prev = bytes(len(rows[0]))
        for r in rows:
            raw += filter_row(r, prev, filters[k % len(filters)], bpp)
            prev = r
            k += 1
    z = zlib.compress(bytes(raw))
    data = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0,
                                                            1 if interlace else 0))
    for t, d in extra:
        data += chunk(t, d)
    step = max(1, (len(z) + idat_parts - 1) // idat_parts)
    for i in range(0, len(z), step):
        data += chunk(b"IDAT", z[i:i + step])
    return data + chunk(b"IEND", b"")


def write(path: str, *args, **kw) -> str:
    with open(path, "wb") as f:
        f.write(encode(*args, **kw))
    return path


def spec_equivalent(kw: dict) -> dict:
    """The encode() arguments whose libpng decode is the expected result of
    `kw`.  libpng 1.6's simplified reader returns rows of the wrong Adam7
    pass for interlaced 16-bit images; the PNG specification fixes the same
    pixels as the non-interlaced file, so that is what is expected."""
    if kw.get("interlace") and kw["depth"] == 16:
        return dict(kw, interlace=False)
    return kw


def random_image(rng: np.random.Generator, h: int, w: int, depth: int, ctype: int,
                 n_palette: int = 0) -> np.ndarray:
    c = CHANNELS[ctype]
    hi = (n_palette if ctype == 3 else (1 << depth))
    return rng.integers(0, hi, size=(h, w, c), dtype=np.int64)


def cases(seed: int = 0) -> List[Tuple[str, dict]]:
    """Deterministic decoding cases: (name, kwargs of encode)."""
    rng = np.random.default_rng(seed)
    out = []
    for ctype, depths in ((0, (1, 2, 4, 8, 16)), (2, (8, 16)), (3, (1, 2, 4, 8)), (4, (8, 16)),
                          (6, (8, 16))):
        for depth in depths:
            for inter in (False, True):
                extra = []
                npal = 0
                if ctype == 3:
                    npal = min(1 << depth, 200)
                    pal = rng.integers(0, 256, size=3 * npal, dtype=np.int64).astype(np.uint8)
                    extra.append((b"PLTE", pal.tobytes()))
                    extra.append((b"tRNS", rng.integers(0, 256, size=max(1, npal // 2),
                                                        dtype=np.int64).astype(np.uint8).tobytes()))
                img = random_image(rng, 13, 11, depth, ctype, npal)
                out.append((f"c{ctype}d{depth}{'i' if inter else ''}",
                            dict(samples=img, depth=depth, ctype=ctype, interlace=inter,
                                 filters=(0, 1, 2, 3, 4, 4, 2), extra=extra, idat_parts=3)))
    # tRNS colour keys
    g = random_image(rng, 9, 10, 8, 0)
    g[0, :3, 0] = 77
    out.append(("key_g8", dict(samples=g, depth=8, ctype=0, extra=[(b"tRNS", struct.pack(">H", 77))])))
    g16 = random_image(rng, 9, 10, 16, 0)
    g16[1, 2, 0] = 40000
    out.append(("key_g16", dict(samples=g16, depth=16, ctype=0,
                                extra=[(b"tRNS", struct.pack(">H", 40000))])))
    g2 = random_image(rng, 5, 9, 2, 0)
    out.append(("key_g2", dict(samples=g2, depth=2, ctype=0, extra=[(b"tRNS", struct.pack(">H", 1))])))
    rgb = random_image(rng, 6, 7, 8, 2)
    rgb[2, 3] = (1, 2, 3)
    out.append(("key_rgb8", dict(samples=rgb, depth=8, ctype=2,
                                 extra=[(b"tRNS", struct.pack(">HHH", 1, 2, 3))])))
    rgb16 = random_image(rng, 6, 7, 16, 2)
    rgb16[0, 0] = (1000, 2000, 3000)
    out.append(("key_rgb16", dict(samples=rgb16, depth=16, ctype=2,
                                  extra=[(b"tRNS", struct.pack(">HHH", 1000, 2000, 3000))])))
    # every 16-bit gray value, linear (no gAMA)
    ramp = np.arange(65536, dtype=np.int64).reshape(256, 256, 1)
    out.append(("ramp16", dict(samples=ramp, depth=16, ctype=0)))
    out.append(("ramp16_srgb", dict(samples=ramp, depth=16, ctype=0, extra=[(b"sRGB", b"\x00")])))
    for sb in (4, 8, 10, 11, 12, 15):
        out.append((f"ramp16_sbit{sb}", dict(samples=ramp, depth=16, ctype=0,
                                             extra=[(b"sBIT", bytes([sb]))])))
    ramp8 = np.arange(256, dtype=np.int64).reshape(16, 16, 1)
    gammas = [20000, 30000, 40000, 43000, 43400, 43500, 43600, 45000, 45455, 47000, 47700,
              47800, 47900, 50000, 60000, 100000, 150000, 220000]
    for gm in gammas:
        out.append((f"ramp8_gama{gm}", dict(samples=ramp8, depth=8, ctype=0,
                                            extra=[(b"gAMA", struct.pack(">I", gm))])))
        out.append((f"ramp16_gama{gm}", dict(samples=ramp, depth=16, ctype=0,
                                             extra=[(b"gAMA", struct.pack(">I", gm))])))
    pal = np.arange(256 * 3, dtype=np.int64) % 256
    pidx = np.arange(256, dtype=np.int64).reshape(16, 16, 1)
    out.append(("pal8_gama100000", dict(samples=pidx, depth=8, ctype=3,
                                        extra=[(b"gAMA", struct.pack(">I", 100000)),
                                               (b"PLTE", pal.astype(np.uint8).tobytes())])))
    ga16 = np.stack([ramp[..., 0], ramp[..., 0][::-1]], -1)
    out.append(("ga16", dict(samples=ga16, depth=16, ctype=4)))
    return out
