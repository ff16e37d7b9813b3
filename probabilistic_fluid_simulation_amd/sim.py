"""The reference's simulation driver (src/main.cpp) on the GPU, in process.

    python -m probabilistic_fluid_simulation_amd.sim <n_timesteps> <delta_t> \\
        <viscosity> <input_image> <velocity_field> [output_dir] [--iters K] ...

runs pfs_sim_main (include/pfs.h) with these arguments; the pfs_sim
executable next to libpfs.so is the same driver as a native program.
"""
from __future__ import annotations

import ctypes
import sys
from typing import Sequence

from . import _lib


def sim_main(args: Sequence[str], prog: str = "pfs_sim") -> int:
    """Runs the driver with argv = [prog, *args]; returns its exit status."""
    argv = [prog, *args]
    arr = (ctypes.c_char_p * (len(argv) + 1))(*[a.encode() for a in argv], None)
    return int(_lib.load().pfs_sim_main(len(argv), arr))


if __name__ == "__main__":
    sys.stdout.flush()
    sys.exit(sim_main(sys.argv[1:], prog="python -m probabilistic_fluid_simulation_amd.sim"))
