"""Particle layer over the grid fluid: tracers and their neighbour grid.

No counterpart in the reference, which has no particles.  Python face of
pfs_particles_advect and pfs_ngrid_* (include/pfs.h): positions are float32
device tensors in field-cell units on the periodic domain [0, W) x [0, H);
integer outputs are int32 tensors holding the C ABI's uint32 values.
"""
from __future__ import annotations

import ctypes
from typing import Tuple

import torch

from . import _lib
from .fluid import VPField, _Bound, _stream


PfsNGrid = _lib.PfsNGrid


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _check_xy(x: torch.Tensor, y: torch.Tensor) -> None:
    if x.dtype != torch.float32 or y.dtype != torch.float32 or x.shape != y.shape or x.dim() != 1:
        raise ValueError("x, y must be 1-D float32 tensors of one length")
    if not (x.is_cuda and y.is_cuda and x.is_contiguous() and y.is_contiguous()):
        raise ValueError("x, y must be contiguous device tensors")


def advect(x: torch.Tensor, y: torch.Tensor, vp: VPField, dt: float, stream=None) -> None:
    """One forward step of the tracers through vp's velocity (in place)."""
    _check_xy(x, y)
    b = _Bound(vp)
    _lib.call("pfs_particles_advect", _ptr(x), _ptr(y), x.numel(), b.ptr(0), dt, _stream(stream))


def scan_exclusive(t: torch.Tensor, stream=None) -> Tuple[torch.Tensor, int]:
    """Exclusive prefix sum of an int32 tensor read as uint32; (out, total)."""
    if t.dtype != torch.int32 or not t.is_contiguous():
        raise ValueError("need a contiguous int32 tensor")
    n = t.numel()
    out = torch.empty_like(t)
    total = torch.zeros(1, dtype=torch.int32, device=t.device)
    ws = torch.empty(((n + 4095) // 4096 + 1,), dtype=torch.int32, device=t.device)
    _lib.call("pfs_scan_exclusive_u32", _ptr(t), _ptr(out), n, _ptr(total), _ptr(ws),
              ws.numel() * 4, _stream(stream))
    return out, int(total.item()) & 0xFFFFFFFF


class NeighborGrid:
    """Uniform grid of cells at least `radius` wide over n particles."""

    def __init__(self, n: int, domain_w: float, domain_h: float, radius: float,
                 device: str = "cuda"):
        self.g = PfsNGrid()
        _lib.call("pfs_ngrid_init", ctypes.byref(self.g), n, domain_w, domain_h, radius)
        i32 = dict(dtype=torch.int32, device=device)
        nc = self.g.gx * self.g.gy
        self.keys = torch.empty(n, **i32)
        self.order = torch.empty(n, **i32)
        self.cell_start = torch.empty(nc, **i32)
        self.cell_end = torch.empty(nc, **i32)
        self.xs = torch.empty(n, dtype=torch.float32, device=device)
        self.ys = torch.empty(n, dtype=torch.float32, device=device)
        nbytes = _lib.load().pfs_ngrid_workspace_bytes(ctypes.byref(self.g))
        self.ws = torch.empty((nbytes + 3) // 4, **i32)
        self.counts = torch.empty(n, **i32)
        self.offsets = torch.empty(n, **i32)
        self.total = torch.zeros(1, **i32)
        self.nbr = None  # list buffer kept between calls of neighbors()

    @property
    def n(self) -> int:
        return self.g.n

    @property
    def cells(self) -> Tuple[int, int]:
        return self.g.gx, self.g.gy

    def build(self, x: torch.Tensor, y: torch.Tensor, stream=None) -> None:
        """Keys, sort order, cell tables and sorted positions for (x, y)."""
        _check_xy(x, y)
        if x.numel() != self.n:
            raise ValueError("particle count differs from the grid's")
        _lib.call("pfs_ngrid_build", ctypes.byref(self.g), _ptr(x), _ptr(y), _ptr(self.keys),
                  _ptr(self.order), _ptr(self.cell_start), _ptr(self.cell_end), _ptr(self.xs),
                  _ptr(self.ys), _ptr(self.ws), self.ws.numel() * 4, _stream(stream))

    def _common(self, xs=None, ys=None, by_position=False):
        xs = self.xs if xs is None else xs
        ys = self.ys if ys is None else ys
        return (ctypes.byref(self.g), _ptr(xs), _ptr(ys), None if by_position else _ptr(self.order),
                _ptr(self.cell_start), _ptr(self.cell_end))

    def neighbor_counts(self, stream=None, xs=None, ys=None, by_position=False):
        """(counts, exclusive offsets) by particle id (by sorted position with
        by_position), without the lists."""
        _lib.call("pfs_ngrid_neighbors", *self._common(xs, ys, by_position), _ptr(self.counts),
                  _ptr(self.offsets), None, 0, _ptr(self.total), _ptr(self.ws),
                  self.ws.numel() * 4, _stream(stream))
        return self.counts, self.offsets

    def neighbors(self, stream=None, xs=None, ys=None, by_position=False, counted=False):
        """(counts, offsets, lists): particle i's neighbours are
        lists[offsets[i] : offsets[i] + counts[i]].  All three are the grid's
        own buffers, overwritten by its next count or list call.  Synchronises
        the stream to size the lists: one counting pass when the list buffer
        kept from an earlier call holds them, else a sizing pass first (the
        buffer then grows with 25% headroom).  counted=True: the grid's counts
        already hold those of these positions (density(..., with_counts=True)
        just ran on them), so no counting pass runs (pfs_ngrid_lists)."""
        fn = "pfs_ngrid_lists" if counted else "pfs_ngrid_neighbors"

        def lists():
            _lib.call(fn, *self._common(xs, ys, by_position), _ptr(self.counts),
                      _ptr(self.offsets), _ptr(self.nbr), self.nbr.numel(), _ptr(self.total),
                      _ptr(self.ws), self.ws.numel() * 4, _stream(stream))
            return self.counts, self.offsets, self.nbr[:int(self.total.item()) & 0xFFFFFFFF]

        if self.nbr is None and counted:  # the first call sizes the buffer from its total
            self.nbr = torch.empty(1, dtype=torch.int32, device=self.keys.device)
        if self.nbr is not None:
            try:
                return lists()
            except _lib.PfsError as e:  # more entries than the buffer: counted, not listed
                if "capacity" not in str(e):
                    raise
        else:
            self.neighbor_counts(stream, xs, ys, by_position)
        total = int(self.total.item()) & 0xFFFFFFFF
        self.nbr = torch.empty(max(total + total // 4, 1), dtype=torch.int32,
                               device=self.keys.device)
        return lists()

    def density(self, mass: float, stream=None, xs=None, ys=None, by_position=False,
                out: torch.Tensor = None, with_counts: bool = False):
        """SPH (2D poly6) density per particle id (per sorted position with
        by_position); with_counts: (density, neighbour counts) from the same
        traversal."""
        rho = out if out is not None else torch.empty(self.n, dtype=torch.float32,
                                                      device=self.keys.device)
        _lib.call("pfs_ngrid_density", *self._common(xs, ys, by_position), mass, _ptr(rho),
                  _ptr(self.counts) if with_counts else None, _ptr(self.ws), self.ws.numel() * 4,
                  _stream(stream))
        return (rho, self.counts) if with_counts else rho


    def pbf(self, mass: float, rho0: float, eps: float, max_iters: int, tol: float = 0.0,
            xs=None, ys=None, stream=None, lists=None):
        """Position-based-fluids density-constraint iterations (pfs_pbf_solve)
        on the sorted positions (xs, ys: those of the last build by default),
        in place, with their neighbour lists found first (which synchronises
        the stream once) unless `lists` brings them.  Returns (lambda by sorted position, residuals: sum
        C^2 per iterate in FP64, -1 where not reached, iterations used) as
        device tensors."""
        xs = self.xs if xs is None else xs
        ys = self.ys if ys is None else ys
        _check_xy(xs, ys)
        lib = _lib.load()
        nbytes = lib.pfs_pbf_workspace_bytes(ctypes.byref(self.g), max_iters)
        if nbytes == 0:
            raise ValueError(f"max_iters must be in [0, {_lib.PFS_PBF_MAX_ITERS}]")
        # lists: (counts, offsets, nbr) of these positions by sorted position,
        # from neighbors(by_position=True), when the caller already has them.
        counts, offsets, nbr = (lists if lists is not None
                                else self.neighbors(stream, xs, ys, by_position=True))
        if nbr.numel() == 0:  # no neighbours at all: the C ABI still wants a buffer
            nbr = torch.zeros(1, dtype=torch.int32, device=xs.device)
        if getattr(self, "_pbf_ws", None) is None or self._pbf_ws.numel() < nbytes + 256:
            self._pbf_ws = torch.empty(nbytes + 256, dtype=torch.uint8, device=xs.device)
        base = (-self._pbf_ws.data_ptr()) % 256
        ws = self._pbf_ws[base:base + nbytes]
        lam = torch.empty(self.n, dtype=torch.float32, device=xs.device)
        res = torch.empty(max_iters + 1, dtype=torch.float64, device=xs.device)
        used = torch.zeros(1, dtype=torch.int32, device=xs.device)
        _lib.call("pfs_pbf_solve", ctypes.byref(self.g), _ptr(xs), _ptr(ys), _ptr(self.cell_start),
                  _ptr(self.cell_end), _ptr(offsets), _ptr(counts), _ptr(nbr), mass, rho0, eps,
                  max_iters, tol, _ptr(lam), _ptr(res), _ptr(used), _ptr(self.ws),
                  self.ws.numel() * 4, _ptr(ws), nbytes, _stream(stream))
        self._pbf_keep = nbr  # the lists stay alive while queued work may read them
        return lam, res, used


def gather(index: torch.Tensor, src: torch.Tensor, out: torch.Tensor, stream=None) -> None:
    """out[p] = src[index[p]] for int32 tensors (uint32 in the C ABI)."""
    for t in (index, src, out):
        if t.dtype != torch.int32 or not t.is_contiguous():
            raise ValueError("need contiguous int32 tensors")
    _lib.call("pfs_gather_u32", _ptr(index), _ptr(src), _ptr(out), index.numel(),
              _stream(stream))


class Tracers:
    """Tracers kept in cell order.  Every sort() rebuilds the neighbour grid
    and adopts its sorted order: positions become the sorted positions and
    `ids` (the original particle numbers) are permuted alongside, so gathers
    from the velocity field, the next sort's input and the neighbour passes
    all touch memory in cell order.  Neighbour lists and densities are then
    indexed by the current (sorted) position."""

    def __init__(self, x: torch.Tensor, y: torch.Tensor, domain_w: float, domain_h: float,
                 radius: float):
        _check_xy(x, y)
        self.x, self.y = x.clone(), y.clone()
        self.ids = torch.arange(x.numel(), dtype=torch.int32, device=x.device)
        self._ids = torch.empty_like(self.ids)
        self.grid = NeighborGrid(x.numel(), domain_w, domain_h, radius, device=str(x.device))

    def advect(self, vp: VPField, dt: float, stream=None) -> None:
        advect(self.x, self.y, vp, dt, stream)

    def sort(self, stream=None) -> None:
        g = self.grid
        g.build(self.x, self.y, stream)
        self.x, g.xs = g.xs, self.x
        self.y, g.ys = g.ys, self.y
        gather(g.order, self.ids, self._ids, stream)
        self.ids, self._ids = self._ids, self.ids

    def neighbor_counts(self, stream=None):
        return self.grid.neighbor_counts(stream, self.x, self.y, by_position=True)

    def neighbors(self, stream=None, counted=False):
        return self.grid.neighbors(stream, self.x, self.y, by_position=True, counted=counted)

    def density(self, mass: float, stream=None, out=None, with_counts: bool = False):
        return self.grid.density(mass, stream, self.x, self.y, by_position=True, out=out,
                                 with_counts=with_counts)

    def pbf(self, mass: float, rho0: float, eps: float, max_iters: int, tol: float = 0.0,
            stream=None):
        """Density-constraint iterations on the tracers' positions (in cell
        order since the last sort), in place; see NeighborGrid.pbf."""
        return self.grid.pbf(mass, rho0, eps, max_iters, tol, self.x, self.y, stream)
