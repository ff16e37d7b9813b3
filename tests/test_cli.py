"""The pfs_sim driver's command line (src/main.cpp:38-41,104-140): argument
errors exit with status 1 and the reference's messages, before any device
work, so these run without a GPU."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from probabilistic_fluid_simulation_amd import _lib, pngio

EXE = os.path.join(_lib.PKG_DIR, "pfs_sim")


def _run(*args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built")
    return subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=60)


@pytest.fixture
def inputs(tmp_path):
    img = np.zeros((4, 5, 4), np.uint8)
    pi, pv = str(tmp_path / "i.png"), str(tmp_path / "v.png")
    pngio.write_png_rgba8(pi, img)
    pngio.write_png_rgba8(pv, img)
    return pi, pv


def test_usage(inputs):
    r = _run(1, 2, 3)
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = _run(1, 0.5, 0.1, *inputs, "out", "extra")
    assert r.returncode == 1 and "Usage:" in r.stderr
    r = _run(1, 0.5, 0.1, *inputs, "--bogus")
    assert r.returncode == 1 and "Usage:" in r.stderr


def test_argument_checks(inputs):
    r = _run(0, 0.5, 0.1, *inputs)
    assert r.returncode == 1 and "Timesteps must be greater than 0." in r.stderr
    r = _run(3, 0, 0.1, *inputs)
    assert r.returncode == 1 and "Delta T must be greater than 0." in r.stderr
    r = _run(3, 0.5, 0.1, *inputs, "--every", "0")
    assert r.returncode == 1


def test_unreadable_inputs(tmp_path, inputs):
    r = _run(3, 0.5, 0.1, str(tmp_path / "missing.png"), inputs[1])
    assert r.returncode == 1 and "Something went wrong reading the input image..." in r.stderr
    r = _run(3, 0.5, 0.1, inputs[0], str(tmp_path / "missing.png"))
    assert r.returncode == 1
    assert "Something went wrong reading the initial velocity field..." in r.stderr


def test_python_entry_point_exports():
    lib = _lib.load()
    assert hasattr(lib, "pfs_sim_main")
