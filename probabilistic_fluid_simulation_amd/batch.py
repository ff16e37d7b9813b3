"""Batched scenes: many independent simulations advanced together.

Python face of the pfs_batch_* C ABI (include/pfs.h).  Each scene is the
reference's four buffers (vp, tmp, image, itmp) as VPFields; one
``step(n)`` advances every scene by n timesteps of the reference driver
(simulate_fluid_step then advect_color_step, src/main.cpp:222-225) with one
launch per kernel stage for all scenes, replayed from a hipGraph.  Results
are bit-identical to stepping each scene alone.
"""
from __future__ import annotations

import ctypes
from typing import Dict, Optional

import torch

from . import _lib
from .fluid import VPField


class Batch:
    def __init__(self, graphs: bool = True):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        # The batch owns a non-blocking stream; callers synchronise with sync().
        _lib.call("pfs_batch_create", ctypes.byref(h), None)
        self._h = h
        self._scenes: Dict[int, tuple] = {}
        self._retired: list = []  # buffers of removed scenes, kept until sync()
        if not graphs:
            _lib.call("pfs_batch_set_graphs", self._h, 0)

    def close(self) -> None:
        if self._h is not None and self._h.value:
            _lib.call("pfs_batch_destroy", self._h)
            self._h = None
            self._scenes.clear()
            self._retired.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, vp: VPField, tmp: VPField, image: VPField, itmp: VPField, dt: float,
            viscosity: float, iters: int, sync: bool = True) -> int:
        """Add a scene (its buffers stay caller-owned); returns its slot.
        sync: first wait for torch's current stream, where the inputs were
        written; a caller that orders the batch's stream after its inputs
        itself (an event) passes False."""
        if sync:
            torch.cuda.current_stream().synchronize()
        slot = ctypes.c_int()
        structs = [f._struct() for f in (vp, tmp, image, itmp)]
        _lib.call("pfs_batch_add", self._h, *[ctypes.byref(s) for s in structs], dt, viscosity,
                  iters, ctypes.byref(slot))
        self._scenes[slot.value] = (vp, tmp, image, itmp)
        return slot.value

    def remove(self, slot: int, hold: bool = True) -> tuple:
        """Remove a scene; later steps no longer touch it.  Its buffers must
        outlive the steps already queued: with hold, the batch keeps them
        until sync(); without, the caller does (they are returned either
        way)."""
        _lib.call("pfs_batch_remove", self._h, slot)
        bufs = self._scenes.pop(slot)
        if hold:
            self._retired.append(bufs)
        return bufs

    def step(self, n: int = 1) -> None:
        _lib.call("pfs_batch_step", self._h, n)

    def sync(self) -> None:
        _lib.call("pfs_batch_sync", self._h)
        self._retired.clear()

    def image(self, slot: int) -> torch.Tensor:
        """The tensor holding the scene's current dye (image and itmp swap
        every step, as in the reference)."""
        im = _lib.PfsField()
        _lib.call("pfs_batch_scene", self._h, slot, ctypes.byref(im), None, None)
        for f in self._scenes[slot][2:]:
            if f.data.data_ptr() == im.data:
                return f.data
        raise RuntimeError("dye buffer not found")

    def buffers(self, slot: int):
        """(vp, tmp, image, itmp) tensors of a scene in the reference's
        post-step roles: image is the current dye."""
        vp, tmp, a, b = self._scenes[slot]
        cur = self.image(slot)
        other = b.data if cur is a.data else a.data
        return vp.data, tmp.data, cur, other

    def steps(self, slot: int) -> int:
        n = ctypes.c_longlong()
        _lib.call("pfs_batch_scene", self._h, slot, None, None, ctypes.byref(n))
        return n.value

    @property
    def count(self) -> int:
        return int(self._lib.pfs_batch_count(self._h))

    @property
    def workspace_bytes(self) -> int:
        return int(self._lib.pfs_batch_workspace_bytes(self._h))

    def set_tile(self, side: int) -> None:
        """Force the fused kernels' tile side (16, 32, 64); 0: automatic."""
        _lib.call("pfs_batch_set_tile", self._h, side)

    @property
    def tile(self) -> int:
        """The tile side the next step uses."""
        side = int(self._lib.pfs_batch_tile(self._h))
        if side < 0:
            raise _lib.PfsError("pfs_batch_tile", side, self._lib.pfs_last_error().decode())
        return side

    def autotune(self, reps: int = 10):
        """Time every tile side on the live scenes (their state is restored)
        and force the fastest; returns (side, {side: ms per step})."""
        ms = (ctypes.c_float * 3)()
        side = ctypes.c_int(0)
        _lib.call("pfs_batch_autotune", self._h, reps, ctypes.byref(side), ms, 3)
        return side.value, {t: float(m) for t, m in zip((64, 32, 16), ms)}

    @property
    def stream_handle(self) -> Optional[int]:
        return self._lib.pfs_batch_stream(self._h)
