"""CPU oracle for the grid fluid step — TEST INFRASTRUCTURE ONLY.

ctypes bindings for two CPU implementations of the reference's solver:

* ``liboracle.so`` (``oracle/pfs_oracle.c``): the plain C restatement of
  ``src/fluid.cpp``, in float (``orc_*``) and a double twin (``orc64_*``).
* ``_ref/libfluidref_k{K}.so``: the reference's own ``src/fluid.cpp``
  compiled from the reference tree by ``oracle/Makefile`` (present only
  where it was built; used to write and re-check the golden vectors and as
  the optional "reference" CPU baseline).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py`` import this
package: ``bench.py`` times it as the CPU baseline and, outside every timed
region, uses it as the checker of its legs' results.  The product library
never touches it.

Fields are numpy arrays of shape (H, W, C), C-contiguous, matching
``vp_field``'s row-major interleaved ``[y][x][c]`` layout
(``includes/fluid.hpp:17-22``).  Functions that swap ``vp_field::data``
pointers in the reference return the arrays in their post-call struct
order, so callers see the same identity changes the reference makes.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")


class _Field32(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int),
                ("data", ctypes.POINTER(ctypes.c_float))]


class _Field64(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int),
                ("data", ctypes.POINTER(ctypes.c_double))]


def _check(a: np.ndarray, dtype) -> None:
    if a.dtype != dtype or a.ndim != 3 or not a.flags["C_CONTIGUOUS"]:
        raise ValueError(f"field must be a C-contiguous (H, W, C) {np.dtype(dtype).name} array")


class _Binder:
    """Maps numpy arrays to vp_field-like structs and back after pointer swaps."""

    def __init__(self, struct, ctype, dtype):
        self.struct, self.ctype, self.dtype = struct, ctype, dtype

    def make(self, *arrays):
        fields = []
        for a in arrays:
            _check(a, self.dtype)
            h, w, c = a.shape
            fields.append(self.struct(w, h, c, a.ctypes.data_as(ctypes.POINTER(self.ctype))))
        return fields

    @staticmethod
    def back(fields, arrays):
        by_addr = {a.ctypes.data: a for a in arrays}
        return tuple(by_addr[ctypes.cast(f.data, ctypes.c_void_p).value] for f in fields)


class Oracle:
    """The C restatement (liboracle.so); ``prec`` 32 or 64 selects the twin."""

    def __init__(self, path: str = ORACLE_SO, prec: int = 32):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        self.lib = ctypes.CDLL(path)
        if prec == 32:
            self.b = _Binder(_Field32, ctypes.c_float, np.float32)
            self.p, self.real = "orc_", ctypes.c_float
        elif prec == 64:
            self.b = _Binder(_Field64, ctypes.c_double, np.float64)
            self.p, self.real = "orc64_", ctypes.c_double
        else:
            raise ValueError("prec must be 32 or 64")
        self.dtype = self.b.dtype

    def _fn(self, name):
        return getattr(self.lib, self.p + name)

    def set_threads(self, n: int) -> None:
        self.lib.orc_set_threads(ctypes.c_int(n))

    def max_threads(self) -> int:
        return int(self.lib.orc_max_threads())

    # --- per-operator entry points (src/fluid.cpp) ---------------------------
    def advect(self, vp, out, dt):
        f = self.b.make(vp, out)
        self._fn("advect")(ctypes.byref(f[0]), ctypes.byref(f[1]), self.real(dt))

    def advect_color(self, image, out, vp, dt):
        f = self.b.make(image, out, vp)
        self._fn("advect_color")(*[ctypes.byref(x) for x in f], self.real(dt))

    def diffuse(self, vp, out, visc, dt, iters) -> Tuple[np.ndarray, np.ndarray]:
        f = self.b.make(vp, out)
        self._fn("diffuse")(ctypes.byref(f[0]), ctypes.byref(f[1]), self.real(visc),
                            self.real(dt), ctypes.c_int(iters))
        return self.b.back(f, (vp, out))

    def add_forces_apply(self, vp, forces, dt):
        """vp[..., :fz] += dt * forces (forces (H, W, fz)), in place."""
        f = self.b.make(vp)
        forces = np.ascontiguousarray(forces, dtype=self.b.dtype)
        if forces.shape[:2] != vp.shape[:2] or forces.ndim != 3:
            raise ValueError("forces must be (H, W, fz)")
        self._fn("add_forces_apply")(ctypes.byref(f[0]),
                                     forces.ctypes.data_as(ctypes.POINTER(self.b.ctype)),
                                     ctypes.c_int(forces.shape[2]), self.real(dt))

    def compute_pressure(self, vp, out, dt, iters) -> Tuple[np.ndarray, np.ndarray]:
        f = self.b.make(vp, out)
        self._fn("compute_pressure")(ctypes.byref(f[0]), ctypes.byref(f[1]), self.real(dt),
                                     ctypes.c_int(iters))
        return self.b.back(f, (vp, out))

    def subtract_pressure_gradient(self, vp, out, dt):
        f = self.b.make(vp, out)
        self._fn("subtract_pressure_gradient")(ctypes.byref(f[0]), ctypes.byref(f[1]),
                                               self.real(dt))

    def simulate_fluid_step(self, vp, tmp, dt, visc, iters):
        f = self.b.make(vp, tmp)
        self._fn("simulate_fluid_step")(ctypes.byref(f[0]), ctypes.byref(f[1]), self.real(dt),
                                        self.real(visc), ctypes.c_int(iters))
        return self.b.back(f, (vp, tmp))

    def advect_color_step(self, image, itmp, vp, dt):
        f = self.b.make(image, itmp, vp)
        self._fn("advect_color_step")(*[ctypes.byref(x) for x in f], self.real(dt))
        return self.b.back(f[:2], (image, itmp))

    def run(self, vp, tmp, image, itmp, dt, visc, iters, steps):
        """``steps`` x (simulate_fluid_step, advect_color_step); returns (vp, tmp, image, itmp)."""
        f = self.b.make(vp, tmp, image, itmp)
        self._fn("run")(*[ctypes.byref(x) for x in f], self.real(dt), self.real(visc),
                        ctypes.c_int(iters), ctypes.c_int(steps))
        return self.b.back(f, (vp, tmp, image, itmp))

    def pressure_residual(self, vp) -> Tuple[float, float]:
        """(sum r^2, sum b^2) in double for p = ch2, b = ch3 of a float field."""
        if self.p != "orc_":
            raise ValueError("residual is defined on float fields")
        f = self.b.make(vp)
        r2, b2 = ctypes.c_double(), ctypes.c_double()
        self.lib.orc_pressure_residual(ctypes.byref(f[0]), ctypes.byref(r2), ctypes.byref(b2))
        return r2.value, b2.value


class RefFluid:
    """The reference's own src/fluid.cpp, built for one Jacobi count K
    (opt "O0": the -O0 build, made for K = 8 only)."""

    def __init__(self, iters: int, opt: str = "O2"):
        suffix = "" if opt == "O2" else f"_{opt}"
        path = os.path.join(REF_DIR, f"libfluidref_k{iters}{suffix}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = ctypes.CDLL(path)
        self.iters = iters
        assert self.lib.ref_jacobi_iters() == iters
        self.b = _Binder(_Field32, ctypes.c_float, np.float32)

    def advect(self, vp, out, dt):
        f = self.b.make(vp, out)
        self.lib.ref_advect(ctypes.byref(f[0]), ctypes.byref(f[1]), ctypes.c_float(dt))

    def advect_color(self, image, out, vp, dt):
        f = self.b.make(image, out, vp)
        self.lib.ref_advect_color(*[ctypes.byref(x) for x in f], ctypes.c_float(dt))

    def diffuse(self, vp, out, visc, dt):
        f = self.b.make(vp, out)
        self.lib.ref_diffuse(ctypes.byref(f[0]), ctypes.byref(f[1]), ctypes.c_float(visc),
                             ctypes.c_float(dt))
        return self.b.back(f, (vp, out))

    def compute_pressure(self, vp, out, dt):
        f = self.b.make(vp, out)
        self.lib.ref_compute_pressure(ctypes.byref(f[0]), ctypes.byref(f[1]), ctypes.c_float(dt))
        return self.b.back(f, (vp, out))

    def subtract_pressure_gradient(self, vp, out, dt):
        f = self.b.make(vp, out)
        self.lib.ref_subtract_pressure_gradient(ctypes.byref(f[0]), ctypes.byref(f[1]),
                                                ctypes.c_float(dt))

    def simulate_fluid_step(self, vp, tmp, dt, visc):
        f = self.b.make(vp, tmp)
        self.lib.ref_simulate_fluid_step(ctypes.byref(f[0]), ctypes.byref(f[1]),
                                         ctypes.c_float(dt), ctypes.c_float(visc))
        return self.b.back(f, (vp, tmp))

    def advect_color_step(self, image, itmp, vp, dt):
        f = self.b.make(image, itmp, vp)
        self.lib.ref_advect_color_step(*[ctypes.byref(x) for x in f], ctypes.c_float(dt))
        return self.b.back(f[:2], (image, itmp))

    def run(self, vp, tmp, image, itmp, dt, visc, steps):
        f = self.b.make(vp, tmp, image, itmp)
        self.lib.ref_run(*[ctypes.byref(x) for x in f], ctypes.c_float(dt), ctypes.c_float(visc),
                         ctypes.c_int(steps))
        return self.b.back(f, (vp, tmp, image, itmp))


def ref_available(iters: int, opt: str = "O2") -> bool:
    suffix = "" if opt == "O2" else f"_{opt}"
    return os.path.exists(os.path.join(REF_DIR, f"libfluidref_k{iters}{suffix}.so"))


_ORACLES = {}


def get(prec: int = 32) -> Oracle:
    if prec not in _ORACLES:
        _ORACLES[prec] = Oracle(prec=prec)
    return _ORACLES[prec]


def ref(iters: int, opt: str = "O2") -> Optional[RefFluid]:
    return RefFluid(iters, opt) if ref_available(iters, opt) else None
