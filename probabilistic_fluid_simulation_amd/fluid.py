"""Python mirror of the reference's solver interface (includes/fluid.hpp).

Same function names and argument meaning as the reference header; the
arithmetic runs in libpfs.so's gfx950 kernels on device tensors, with
results identical to the reference's CPU solver src/fluid.cpp.

``VPField`` plays the role of ``vp_field`` (includes/fluid.hpp:17-22): a
width x height x channels float32 field in row-major interleaved
``[y][x][c]`` order, here a contiguous CUDA tensor of shape (H, W, C).
Functions that swap ``vp_field::data`` pointers in the reference
(``diffuse``, ``computePressure``, ``advect_color_step``) swap the
``data`` tensors of the VPField objects passed in, exactly as the
reference does.

Work is enqueued on torch's current stream and returns without waiting.
Errors raise ``PfsError`` (the reference's functions return void and
check nothing).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from . import _lib

NUM_JACOBI_ITERS = 30  # includes/fluid.hpp:11; every function below takes it as `iters`


class VPField:
    """vp_field: x = width, y = height, z = channels, data = (H, W, C) float32 tensor."""

    __slots__ = ("data",)

    def __init__(self, data: torch.Tensor):
        if data.dtype != torch.float32 or data.dim() != 3:
            raise ValueError("VPField data must be a 3-D float32 tensor (H, W, C)")
        if not data.is_cuda or not data.is_contiguous():
            raise ValueError("VPField data must be a contiguous CUDA tensor")
        self.data = data

    @property
    def x(self) -> int:
        return self.data.shape[1]

    @property
    def y(self) -> int:
        return self.data.shape[0]

    @property
    def z(self) -> int:
        return self.data.shape[2]

    @classmethod
    def from_numpy(cls, a: np.ndarray, device: str = "cuda") -> "VPField":
        return cls(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device))

    @classmethod
    def empty(cls, w: int, h: int, c: int = 4, device: str = "cuda") -> "VPField":
        return cls(torch.empty((h, w, c), dtype=torch.float32, device=device))

    def numpy(self) -> np.ndarray:
        return self.data.detach().cpu().numpy()

    def _struct(self) -> _lib.PfsField:
        return _lib.PfsField(self.x, self.y, self.z, self.data.data_ptr())


def _stream(stream: Optional[torch.cuda.Stream]) -> ctypes.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class _Bound:
    """pfs_field structs for a set of VPFields; writes swapped pointers back."""

    def __init__(self, *fields: VPField):
        self.fields = fields
        self.structs = [f._struct() for f in fields]
        self.tensors = {f.data.data_ptr(): f.data for f in fields}

    def ptr(self, k: int):
        return ctypes.byref(self.structs[k])

    def sync_back(self) -> None:
        for f, s in zip(self.fields, self.structs):
            f.data = self.tensors[s.data]


# --- the operators of includes/fluid.hpp ---------------------------------------
def advect(vp: VPField, vp_out: VPField, dt: float, stream=None) -> None:
    """includes/fluid.hpp:32, src/fluid.cpp:24-70."""
    b = _Bound(vp, vp_out)
    _lib.call("pfs_advect", b.ptr(0), b.ptr(1), dt, _stream(stream))


def advect_color(image: VPField, out: VPField, vp: VPField, dt: float, stream=None) -> None:
    """includes/fluid.hpp:45, src/fluid.cpp:72-127."""
    b = _Bound(image, out, vp)
    _lib.call("pfs_advect_color", b.ptr(0), b.ptr(1), b.ptr(2), dt, _stream(stream))


def diffuse(vp: VPField, vp_out: VPField, viscosity: float, dt: float,
            iters: int = NUM_JACOBI_ITERS, stream=None) -> None:
    """includes/fluid.hpp:58, src/fluid.cpp:129-196 (swaps vp.data/vp_out.data
    after every sweep but the last)."""
    b = _Bound(vp, vp_out)
    _lib.call("pfs_diffuse", b.ptr(0), b.ptr(1), viscosity, dt, iters, _stream(stream))
    b.sync_back()


def addForces(vp: VPField, forces=None, stream=None) -> None:  # noqa: N802 (reference name)
    """includes/fluid.hpp:70, src/fluid.cpp:198-208: a no-op."""
    b = _Bound(vp)
    _lib.call("pfs_add_forces", b.ptr(0), None, _stream(stream))


def add_forces_apply(vp: VPField, forces: torch.Tensor, dt: float, stream=None) -> None:
    """Force injection the reference leaves as a stub (src/fluid.cpp:198-208,
    its call commented out at :302): vp[..., :fz] += dt * forces, forces an
    (H, W, fz) float32 device tensor, fz <= channels."""
    if forces.dtype != torch.float32 or forces.dim() != 3 or not forces.is_contiguous():
        raise ValueError("forces must be a contiguous (H, W, fz) float32 tensor")
    if tuple(forces.shape[:2]) != (vp.y, vp.x):
        raise ValueError("forces must have the field's height and width")
    b = _Bound(vp)
    _lib.call("pfs_add_forces_apply", b.ptr(0), ctypes.c_void_p(forces.data_ptr()),
              int(forces.shape[2]), dt, _stream(stream))


def computePressure(vp: VPField, vp_out: VPField, dt: float, iters: int = NUM_JACOBI_ITERS,
                    stream=None) -> None:  # noqa: N802
    """includes/fluid.hpp:81, src/fluid.cpp:210-267."""
    b = _Bound(vp, vp_out)
    _lib.call("pfs_compute_pressure", b.ptr(0), b.ptr(1), dt, iters, _stream(stream))
    b.sync_back()


def subtractPressureGradient(vp: VPField, vp_out: VPField, dt: float,
                             stream=None) -> None:  # noqa: N802
    """includes/fluid.hpp:92, src/fluid.cpp:269-296."""
    b = _Bound(vp, vp_out)
    _lib.call("pfs_subtract_pressure_gradient", b.ptr(0), b.ptr(1), dt, _stream(stream))


def simulate_fluid_step(vp: VPField, tmp: VPField, dt: float, viscosity: float,
                        iters: int = NUM_JACOBI_ITERS, stream=None) -> None:
    """includes/fluid.hpp:109, src/fluid.cpp:298-310."""
    b = _Bound(vp, tmp)
    _lib.call("pfs_simulate_fluid_step", b.ptr(0), b.ptr(1), dt, viscosity, iters,
              _stream(stream))
    b.sync_back()


def advect_color_step(image: VPField, itmp: VPField, vp: VPField, dt: float,
                      stream=None) -> None:
    """includes/fluid.hpp:118, src/fluid.cpp:312-320 (swaps image.data/itmp.data)."""
    b = _Bound(image, itmp, vp)
    _lib.call("pfs_advect_color_step", b.ptr(0), b.ptr(1), b.ptr(2), dt, _stream(stream))
    b.sync_back()


def fluid_timestep(vp: VPField, tmp: VPField, image: VPField, itmp: VPField, dt: float,
                   viscosity: float, iters: int = NUM_JACOBI_ITERS, stream=None) -> None:
    """One timestep of the reference driver loop (src/main.cpp:222-225):
    simulate_fluid_step then advect_color_step, in one call so that the dye
    advection can run inside the step's last kernel.  Swaps image.data and
    itmp.data like advect_color_step."""
    b = _Bound(vp, tmp, image, itmp)
    _lib.call("pfs_fluid_timestep", b.ptr(0), b.ptr(1), b.ptr(2), b.ptr(3), dt, viscosity, iters,
              _stream(stream))
    b.sync_back()


def fluid_timestep_noalloc(vp: VPField, tmp: VPField, image: VPField, itmp: VPField, dt: float,
                           viscosity: float, iters: int = NUM_JACOBI_ITERS, stream=None) -> None:
    """fluid_timestep with the step's workspace in the library's per-GPU
    device array (pfs_fluid_timestep_noalloc): allocates nothing; fields
    whose workspace exceeds PFS_NOALLOC_SCRATCH_BYTES are rejected."""
    b = _Bound(vp, tmp, image, itmp)
    _lib.call("pfs_fluid_timestep_noalloc", b.ptr(0), b.ptr(1), b.ptr(2), b.ptr(3), dt,
              viscosity, iters, _stream(stream))
    b.sync_back()


def fluid_run(vp: VPField, tmp: VPField, image: VPField, itmp: VPField, dt: float,
              viscosity: float, iters: int, nsteps: int, stream=None) -> None:
    """nsteps timesteps in one call (fluid_timestep nsteps times, identical
    results), replayed from hipGraphs of 16 and 2 steps kept for these
    buffers.  Swaps image.data and itmp.data nsteps times."""
    b = _Bound(vp, tmp, image, itmp)
    _lib.call("pfs_fluid_run", b.ptr(0), b.ptr(1), b.ptr(2), b.ptr(3), dt, viscosity, iters,
              nsteps, _stream(stream))
    b.sync_back()


_precise_ws = {}


def _workspace(key, nbytes: int, device) -> torch.Tensor:
    ws = _precise_ws.get(key)
    if ws is None or ws.numel() < nbytes or ws.device != device:
        ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
        _precise_ws[key] = ws
    return ws


def simulate_fluid_step_precise(vp: VPField, tmp: VPField, dt: float, viscosity: float,
                                max_iters: int = 16, tol: float = 1e-4, stream=None,
                                report: bool = True) -> Optional[dict]:
    """The precision form (no reference counterpart): simulate_fluid_step
    with compensated stencil sums, an FP64 pressure residual per sweep and
    early exit at ||r|| <= tol ||b|| (at most max_iters sweeps).  Returns the
    report, which synchronises the stream; report=False queues the step
    without waiting and returns None."""
    lib = _lib.load()
    nbytes = lib.pfs_precise_workspace_bytes(vp.x, vp.y, 0)
    ws = _workspace(("f32", vp.data.device), nbytes, vp.data.device)
    rep = _lib.PreciseReport()
    b = _Bound(vp, tmp)
    _lib.call("pfs_simulate_fluid_step_precise", b.ptr(0), b.ptr(1), dt, viscosity, max_iters,
              tol, ctypes.c_void_p(ws.data_ptr()), ws.numel(), _stream(stream),
              ctypes.byref(rep) if report else None)
    if not report:
        return None
    return {"iters_used": rep.iters_used, "b_norm": rep.b_norm,
            "residual_norm": rep.residual_norm,
            "residual_history": list(rep.residual_history)[:max(rep.iters_used + 2, 1)]}


def set_precise_schedule(sweeps_per_launch: int = 0, tile: int = 0,
                         pressure_threads: int = 0) -> None:
    """The precision form's sweeps per launch (4 or 8), tile side (16, 32,
    64) and pressure workgroup threads at side 64 (512 or 1024); 0: the
    default of each.  The outputs do not depend on the schedule."""
    _lib.call("pfs_set_precise_schedule", sweeps_per_launch, tile, pressure_threads)


def simulate_fluid_step_f64(vp: torch.Tensor, tmp: torch.Tensor, dt: float, viscosity: float,
                            iters_diffuse: int, iters_pressure: int, stream=None) -> None:
    """FP64 twin of the step on (H, W, 4) float64 device tensors: plain double
    arithmetic, given sweep counts, the precision form's buffer contract."""
    if vp.dtype != torch.float64 or tmp.dtype != torch.float64 or vp.shape != tmp.shape:
        raise ValueError("vp and tmp must be float64 tensors of one shape (H, W, 4)")
    h, w, c = vp.shape
    if c != 4 or not (vp.is_contiguous() and tmp.is_contiguous()):
        raise ValueError("need contiguous (H, W, 4) tensors")
    lib = _lib.load()
    nbytes = lib.pfs_precise_workspace_bytes(w, h, 1)
    ws = _workspace(("f64", vp.device), nbytes, vp.device)
    _lib.call("pfs_simulate_fluid_step_f64", ctypes.c_void_p(vp.data_ptr()),
              ctypes.c_void_p(tmp.data_ptr()), w, h, dt, viscosity, iters_diffuse,
              iters_pressure, ctypes.c_void_p(ws.data_ptr()), ws.numel(), _stream(stream))


# --- library controls -------------------------------------------------------------
def set_step_path(path: str) -> None:
    """"auto" (fused gfx950 kernels) or "per_op" (one launch per operator/sweep)."""
    code = {"auto": _lib.PFS_PATH_AUTO, "per_op": _lib.PFS_PATH_PER_OP}[path]
    _lib.call("pfs_set_step_path", code)


def get_step_path() -> str:
    return {_lib.PFS_PATH_AUTO: "auto", _lib.PFS_PATH_PER_OP: "per_op"}[
        _lib.load().pfs_get_step_path()]


TILE_SIDES = (64, 32, 16)


def set_step_tile(side: int) -> None:
    """Force the fused step's tile side (16, 32 or 64); 0 returns to the
    automatic choice (the autotuner's record for the shape, else the rule)."""
    _lib.call("pfs_set_step_tile", side)


def get_step_tile(w: int, h: int, iters: int, dye_fused: bool = False) -> int:
    """The tile side a step of this shape takes now."""
    lib = _lib.load()
    side = lib.pfs_get_step_tile(w, h, iters, int(bool(dye_fused)))
    if side < 0:
        raise _lib.PfsError("pfs_get_step_tile", side, lib.pfs_last_error().decode(errors="replace"))
    return side


def autotune_step(vp: VPField, tmp: VPField, dt: float, viscosity: float, iters: int,
                  image: Optional[VPField] = None, itmp: Optional[VPField] = None,
                  reps: int = 20, stream=None):
    """Time the step (with image/itmp: the whole fluid_timestep) on scratch
    copies for every tile side and record the fastest for this shape; the
    buffers are not changed.  Returns (side, {side: ms per step})."""
    ms = (ctypes.c_float * 3)()
    side = ctypes.c_int(0)
    if image is None:
        b = _Bound(vp, tmp)
        args = (b.ptr(0), b.ptr(1), None, None)
    else:
        b = _Bound(vp, tmp, image, itmp)
        args = (b.ptr(0), b.ptr(1), b.ptr(2), b.ptr(3))
    _lib.call("pfs_autotune_step", *args, dt, viscosity, iters, reps, _stream(stream),
              ctypes.byref(side), ms, 3)
    return side.value, {t: float(m) for t, m in zip(TILE_SIDES, ms)}


def clear_step_tuning() -> None:
    """Forget every side autotune_step recorded."""
    _lib.call("pfs_clear_step_tuning")


def profile_step(vp: VPField, tmp: VPField, dt: float, viscosity: float, iters: int,
                 image: Optional[VPField] = None, itmp: Optional[VPField] = None, stream=None,
                 reps: int = 1):
    """One fused step (or, with image/itmp, one fluid_timestep) timed per
    kernel group with HIP events on the launch stream, each group launched
    `reps` times back to back (the buffers still advance by one step):
    returns (advect+diffusion+divergence ms, pressure+projection(+dye) ms)
    per launch."""
    ms = (ctypes.c_float * 2)()
    if image is None:
        b = _Bound(vp, tmp)
        args = (b.ptr(0), b.ptr(1), None, None)
    else:
        b = _Bound(vp, tmp, image, itmp)
        args = (b.ptr(0), b.ptr(1), b.ptr(2), b.ptr(3))
    _lib.call("pfs_profile_step", *args, dt, viscosity, iters, reps, _stream(stream), ms, 2)
    b.sync_back()
    return float(ms[0]), float(ms[1])


def advect_keys(vp: VPField, dt: float, stream=None) -> torch.Tensor:
    """Back-trace cell keys (i0, j0, i1, j1) of advect for every cell, (H, W, 4) int32."""
    keys = torch.empty((vp.y, vp.x, 4), dtype=torch.int32, device=vp.data.device)
    b = _Bound(vp)
    _lib.call("pfs_debug_advect_keys", b.ptr(0), dt, ctypes.c_void_p(keys.data_ptr()),
              _stream(stream))
    return keys


def advect_color_keys(image: VPField, vp: VPField, dt: float, stream=None) -> torch.Tensor:
    """(vi, vj, i0, j0, i1, j1) of advect_color for every pixel, (Hi, Wi, 6) int32."""
    keys = torch.empty((image.y, image.x, 6), dtype=torch.int32, device=vp.data.device)
    b = _Bound(image, vp)
    _lib.call("pfs_debug_advect_color_keys", b.ptr(0), b.ptr(1), dt,
              ctypes.c_void_p(keys.data_ptr()), _stream(stream))
    return keys
