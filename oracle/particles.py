"""ctypes binding of oracle/pfs_particles.c, the CPU restatement of the
particle layer -- TEST INFRASTRUCTURE ONLY (see that file's header)."""
from __future__ import annotations

import ctypes
from typing import Tuple

import numpy as np

from . import ORACLE_SO

_lib = None


def _load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(ORACLE_SO)
        P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
        lib.orcp_advect.argtypes = [P, P, I, P, I, I, I, F]
        lib.orcp_advect.restype = None
        lib.orcp_ngrid_build.argtypes = [P, P, I, F, F, I, I, P, P, P, P, P, P]
        lib.orcp_ngrid_build.restype = I
        lib.orcp_ngrid_neighbors.argtypes = [P, P, P, I, P, P, P, I, I, F, F, F, P, P, P]
        lib.orcp_ngrid_neighbors.restype = ctypes.c_longlong
        lib.orcp_ngrid_density.argtypes = [P, P, P, I, P, P, P, I, I, F, F, F, F, P]
        lib.orcp_ngrid_density.restype = None
        lib.orcp_pbf_solve.argtypes = [P, P, I, P, P, P, F, F, F, F, F, F, I, ctypes.c_double,
                                       P, P, ctypes.POINTER(I)]
        lib.orcp_pbf_solve.restype = I
        _lib = lib
    return _lib


def _p(a: np.ndarray) -> ctypes.c_void_p:
    assert a.flags.c_contiguous
    return ctypes.c_void_p(a.ctypes.data)


def advect(x: np.ndarray, y: np.ndarray, vp: np.ndarray, dt: float) -> None:
    """In place, float32 x, y; vp (H, W, C) float32."""
    h, w, c = vp.shape
    vp = np.ascontiguousarray(vp, dtype=np.float32)
    _load().orcp_advect(_p(x), _p(y), len(x), _p(vp), w, h, c, dt)


class Grid:
    """The neighbour grid of n particles (parameters as pfs_ngrid)."""

    def __init__(self, x, y, W, H, r, gx, gy, kx, ky):
        self.x = np.ascontiguousarray(x, dtype=np.float32)
        self.y = np.ascontiguousarray(y, dtype=np.float32)
        self.n = len(self.x)
        self.W, self.H, self.r, self.gx, self.gy = W, H, r, gx, gy
        nc = gx * gy
        self.keys = np.empty(self.n, np.uint32)
        self.order = np.empty(self.n, np.uint32)
        self.start = np.empty(nc, np.uint32)
        self.end = np.empty(nc, np.uint32)
        self.xs = np.empty(self.n, np.float32)
        self.ys = np.empty(self.n, np.float32)
        rc = _load().orcp_ngrid_build(_p(self.x), _p(self.y), self.n, kx, ky, gx, gy,
                                      _p(self.keys), _p(self.order), _p(self.start),
                                      _p(self.end), _p(self.xs), _p(self.ys))
        assert rc == 0
        self.keys_sorted = np.ascontiguousarray(self.keys[self.order])

    def neighbors(self, by_position: bool = False) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(counts by id, exclusive offsets, lists); by sorted position with
        by_position."""
        lib = _load()
        counts = np.empty(self.n, np.uint32)
        order = None if by_position else _p(self.order)
        args = (_p(self.xs), _p(self.ys), order, self.n, _p(self.keys_sorted),
                _p(self.start), _p(self.end), self.gx, self.gy, self.W, self.H, self.r)
        total = lib.orcp_ngrid_neighbors(*args, _p(counts), None, None)
        offsets = np.zeros(self.n, np.uint32)
        offsets[1:] = np.cumsum(counts, dtype=np.uint64)[:-1].astype(np.uint32)
        nbr = np.empty(max(int(total), 1), np.uint32)
        lib.orcp_ngrid_neighbors(*args, _p(counts), _p(offsets), _p(nbr))
        return counts, offsets, nbr[:int(total)]

    def density(self, mass: float, by_position: bool = False) -> np.ndarray:
        rho = np.empty(self.n, np.float32)
        order = None if by_position else _p(self.order)
        _load().orcp_ngrid_density(_p(self.xs), _p(self.ys), order, self.n,
                                   _p(self.keys_sorted), _p(self.start), _p(self.end), self.gx,
                                   self.gy, self.W, self.H, self.r, mass, _p(rho))
        return rho

    def pbf(self, mass: float, rho0: float, eps: float, max_iters: int, tol: float = 0.0):
        """The position-based-fluids solve on copies of the sorted positions,
        with their neighbour lists: (xs, ys, lambda of the last iterate,
        residuals [max_iters + 1], iters used)."""
        counts, offsets, nbr = self.neighbors(by_position=True)
        nbr = np.ascontiguousarray(nbr if len(nbr) else np.zeros(1, np.uint32))
        xs, ys = self.xs.copy(), self.ys.copy()
        lam = np.empty(self.n, np.float32)
        res = np.empty(max_iters + 1, np.float64)
        used = ctypes.c_int(0)
        rc = _load().orcp_pbf_solve(_p(xs), _p(ys), self.n, _p(offsets), _p(counts), _p(nbr),
                                    self.W, self.H, self.r, mass, rho0, eps, max_iters, tol,
                                    _p(lam), _p(res), ctypes.byref(used))
        assert rc == 0
        return xs, ys, lam, res, used.value
