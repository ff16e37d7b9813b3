"""ctypes binding of libpfs.so (the C ABI declared in include/pfs.h).

The shared library is built in-tree by ``make -C probabilistic_fluid_simulation_amd/csrc``
(or ``__graft_entry__.build()``).  There is no fallback: if the library is
missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libpfs.so")

PFS_OK = 0
PFS_ERR_INVALID = -1
PFS_ERR_HIP = -2
PFS_ERR_NOMEM = -3
PFS_ERR_UNSUPPORTED = -4
PFS_PATH_AUTO = 0
PFS_PATH_PER_OP = 1
PFS_PRECISE_MAX_ITERS = 64
PFS_PBF_MAX_ITERS = 64

_ERRNAMES = {PFS_ERR_INVALID: "invalid argument", PFS_ERR_HIP: "HIP error",
             PFS_ERR_NOMEM: "out of device memory", PFS_ERR_UNSUPPORTED: "unsupported"}


class PfsField(ctypes.Structure):
    """pfs_field / vp_field: {int x, y, z; float *data} (includes/fluid.hpp:17-22)."""
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int),
                ("data", ctypes.c_void_p)]


class PreciseReport(ctypes.Structure):
    """pfs_precise_report (include/pfs.h)."""
    _fields_ = [("iters_used", ctypes.c_int), ("b_norm", ctypes.c_double),
                ("residual_norm", ctypes.c_double),
                ("residual_history", ctypes.c_double * (PFS_PRECISE_MAX_ITERS + 1))]


class PfsError(RuntimeError):
    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn}: {_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class PfsNGrid(ctypes.Structure):
    """pfs_ngrid (include/pfs.h): the particle layer's neighbour grid."""
    _fields_ = [("n", ctypes.c_int), ("domain_w", ctypes.c_float), ("domain_h", ctypes.c_float),
                ("radius", ctypes.c_float), ("gx", ctypes.c_int), ("gy", ctypes.c_int),
                ("kx", ctypes.c_float), ("ky", ctypes.c_float)]


_P = ctypes.POINTER(PfsField)
_G = ctypes.POINTER(PfsNGrid)
_V = ctypes.c_void_p
_F = ctypes.c_float
_I = ctypes.c_int
_FPP = ctypes.POINTER(ctypes.c_void_p)

# name -> argtypes (every function returns int status unless listed in _RESTYPES)
SIGNATURES = {
    "pfs_advect": [_P, _P, _F, _V],
    "pfs_advect_color": [_P, _P, _P, _F, _V],
    "pfs_diffuse": [_P, _P, _F, _F, _I, _V],
    "pfs_add_forces": [_P, _V, _V],
    "pfs_add_forces_apply": [_P, _V, _I, _F, _V],
    "pfs_compute_pressure": [_P, _P, _F, _I, _V],
    "pfs_subtract_pressure_gradient": [_P, _P, _F, _V],
    "pfs_simulate_fluid_step": [_P, _P, _F, _F, _I, _V],
    "pfs_advect_color_step": [_P, _P, _P, _F, _V],
    "pfs_fluid_timestep": [_P, _P, _P, _P, _F, _F, _I, _V],
    "pfs_fluid_run": [_P, _P, _P, _P, _F, _F, _I, _I, _V],
    "pfs_fluid_timestep_noalloc": [_P, _P, _P, _P, _F, _F, _I, _V],
    "pfs_simulate_fluid_step_dev": [_FPP, _FPP, _F, _F, _I, _I, _I, _I, _V],
    "pfs_advect_color_step_dev": [_FPP, _FPP, _FPP, _F, _I, _I, _I, _I, _I, _I, _V],
    "pfs_set_step_path": [_I],
    "pfs_set_step_tile": [_I],
    "pfs_get_step_tile": [_I, _I, _I, _I],
    "pfs_autotune_step": [_P, _P, _P, _P, _F, _F, _I, _I, _V, ctypes.POINTER(ctypes.c_int),
                          ctypes.POINTER(ctypes.c_float), _I],
    "pfs_clear_step_tuning": [],
    "pfs_get_step_path": [],
    "pfs_step_workspace_bytes": [_I, _I],
    "pfs_profile_step": [_P, _P, _P, _P, _F, _F, _I, _I, _V, ctypes.POINTER(ctypes.c_float), _I],
    "pfs_debug_advect_keys": [_P, _F, _V, _V],
    "pfs_debug_advect_color_keys": [_P, _P, _F, _V, _V],
    "pfs_batch_create": [ctypes.POINTER(ctypes.c_void_p), _V],
    "pfs_batch_destroy": [_V],
    "pfs_batch_add": [_V, _P, _P, _P, _P, _F, _F, _I, ctypes.POINTER(ctypes.c_int)],
    "pfs_batch_remove": [_V, _I],
    "pfs_batch_step": [_V, _I],
    "pfs_batch_sync": [_V],
    "pfs_batch_scene": [_V, _I, _P, _P, ctypes.POINTER(ctypes.c_longlong)],
    "pfs_batch_count": [_V],
    "pfs_batch_workspace_bytes": [_V],
    "pfs_batch_set_graphs": [_V, _I],
    "pfs_batch_stream": [_V],
    "pfs_batch_set_tile": [_V, _I],
    "pfs_batch_tile": [_V],
    "pfs_batch_autotune": [_V, _I, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float),
                           _I],
    "pfs_precise_workspace_bytes": [_I, _I, _I],
    "pfs_set_precise_schedule": [_I, _I, _I],
    "pfs_simulate_fluid_step_precise": [_P, _P, _F, _F, _I, ctypes.c_double, _V, ctypes.c_size_t,
                                        _V, ctypes.POINTER(PreciseReport)],
    "pfs_simulate_fluid_step_f64": [_V, _V, _I, _I, ctypes.c_double, ctypes.c_double, _I, _I, _V,
                                    ctypes.c_size_t, _V],
    "pfs_png_size": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
    "pfs_png_read_rgba8": [ctypes.c_char_p, _V, ctypes.c_size_t],
    "pfs_png_read": [ctypes.c_char_p, _V, ctypes.c_size_t],
    "pfs_png_write_rgba8": [ctypes.c_char_p, _V, _I, _I],
    "pfs_png_write": [ctypes.c_char_p, _V, _I, _I],
    "pfs_sim_main": [_I, ctypes.POINTER(ctypes.c_char_p)],
    "pfs_particles_advect": [_V, _V, _I, _P, _F, _V],
    "pfs_ngrid_init": [_G, _I, _F, _F, _F],
    "pfs_ngrid_workspace_bytes": [_G],
    "pfs_pbf_workspace_bytes": [_G, _I],
    "pfs_pbf_solve": [_G, _V, _V, _V, _V, _V, _V, _V, _F, _F, _F, _I, ctypes.c_double, _V, _V,
                      _V, _V, ctypes.c_size_t, _V, ctypes.c_size_t, _V],
    "pfs_ngrid_build": [_G, _V, _V, _V, _V, _V, _V, _V, _V, _V, ctypes.c_size_t, _V],
    "pfs_ngrid_neighbors": [_G, _V, _V, _V, _V, _V, _V, _V, _V, ctypes.c_size_t, _V, _V,
                            ctypes.c_size_t, _V],
    "pfs_ngrid_lists": [_G, _V, _V, _V, _V, _V, _V, _V, _V, ctypes.c_size_t, _V, _V,
                        ctypes.c_size_t, _V],
    "pfs_ngrid_density": [_G, _V, _V, _V, _V, _V, _F, _V, _V, _V, ctypes.c_size_t, _V],
    "pfs_scan_exclusive_u32": [_V, _V, _I, _V, _V, ctypes.c_size_t, _V],
    "pfs_gather_u32": [_V, _V, _V, _I, _V],
    "pfs_version": [],
    "pfs_last_error": [],
}
_RESTYPES = {"pfs_version": ctypes.c_char_p, "pfs_last_error": ctypes.c_char_p,
             "pfs_step_workspace_bytes": ctypes.c_size_t,
             "pfs_batch_workspace_bytes": ctypes.c_size_t, "pfs_batch_stream": ctypes.c_void_p,
             "pfs_precise_workspace_bytes": ctypes.c_size_t,
             "pfs_ngrid_workspace_bytes": ctypes.c_size_t,
             "pfs_pbf_workspace_bytes": ctypes.c_size_t}

_lock = threading.Lock()
_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load libpfs.so once; raises if it was not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise ImportError(f"{path} is not built: run `make -C {PKG_DIR}/csrc` or "
                                  "__graft_entry__.build()")
            lib = ctypes.CDLL(path)
            for name, args in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = _RESTYPES.get(name, ctypes.c_int)
            _lib = lib
    return _lib


def call(name: str, *args) -> int:
    """Call a status-returning entry point; raise PfsError on failure."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != PFS_OK:
        raise PfsError(name, rc, lib.pfs_last_error().decode(errors="replace"))
    return rc


def version() -> str:
    return load().pfs_version().decode()
