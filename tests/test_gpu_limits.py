"""The largest fields the library's 32-bit indexing accepts (x * y * z below
2^31 floats; one more cell is rejected, tests/test_abi.py).  One timestep,
through the fused kernels and through the reference's two calls, on inputs
where src/fluid.cpp's step is an exact fixed point: zero velocity, pressure
1, divergence -0 (what gamma * (+0) gives, gamma = -1/dt), and any finite
positive dye.  Then every bit of vp, tmp and the dye must come back:

  advect      x' = i - (dt*0)/W = i, corner weights (1, 0, 0, 0): u = 0
  diffuse     (a*0 + a*0 + a*0 + a*0 + 1*0) / beta = +0
  divergence  gamma * ((0 - 0) + (0 - 0)) = -0
  pressure    (1 + 1 + 1 + 1 + 1*(-0)) / 4 = 1
  gradient    ((1 - 1)*dt)/2 = 0, u - 0 = +0
  dye         (1-0)(1-0) f00 + 0 f10 + 0 f01 + 0 f11 = f00

so a cell or channel offset that overflowed, a tile that was skipped or
written twice, or a workspace plane that wrapped would change some bits.
The library keeps the step's workspace for the stream it ran on (32 B per
cell here, about 17 GB) until the process ends."""
from __future__ import annotations

import pytest

torch = pytest.importorskip("torch")

from probabilistic_fluid_simulation_amd import fluid  # noqa: E402

pytestmark = pytest.mark.gpu

SHAPES = [(23170, 23170), (65535, 8192)]  # w * h * 4 <= 2^31 - 1; one more column is not


def _fixed_point(w, h):
    vp = torch.zeros((h, w, 4), device="cuda")
    vp[..., 2] = 1.0
    vp[..., 3] = -0.0
    x = torch.arange(w, device="cuda", dtype=torch.int32).view(1, w)
    y = torch.arange(h, device="cuda", dtype=torch.int32).view(h, 1)
    img = torch.empty((h, w, 4), device="cuda")
    for c in range(4):  # a finite, positive pattern that differs between neighbours
        img[..., c] = ((x * 7 + y * 13 + 3 * c) % 251 + 1).float() / 256.0
    return vp, img


def _bits(t):
    return t.view(torch.int32)


@pytest.mark.parametrize("w,h", SHAPES, ids=[f"{w}x{h}" for w, h in SHAPES])
def test_largest_fields_fixed_point(w, h):
    assert w * h * 4 <= 2**31 - 1 < (w + 1) * h * 4
    # vp, tmp, image, itmp, copies of vp and image, the step's workspace
    # (32 B per cell) and the pattern's temporaries: about 10 buffers' worth.
    need = 10 * w * h * 16
    free, _ = torch.cuda.mem_get_info()
    if free < need:
        pytest.skip(f"needs {need / 2**30:.0f} GiB of free device memory, {free / 2**30:.0f} free")
    vp, img = _fixed_point(w, h)
    vp0, img0 = vp.clone(), img.clone()
    dt, visc, k = 1.0 / 120.0, 0.1, 8

    # the fused timestep: dye advected inside the pressure kernel
    V, T = fluid.VPField(vp), fluid.VPField(vp.clone())
    I, IT = fluid.VPField(img), fluid.VPField(torch.empty_like(img))
    fluid.fluid_timestep(V, T, I, IT, dt, visc, k)
    torch.cuda.synchronize()
    assert torch.equal(_bits(V.data), _bits(vp0)), "vp after the fused step"
    assert torch.equal(_bits(T.data), _bits(vp0)), "tmp after the fused step"
    assert torch.equal(_bits(I.data), _bits(img0)), "dye after the fused step"

    # the reference's two calls: simulate_fluid_step, then advect_color_step
    fluid.simulate_fluid_step(V, T, dt, visc, k)
    fluid.advect_color_step(I, IT, V, dt)
    torch.cuda.synchronize()
    assert torch.equal(_bits(V.data), _bits(vp0)), "vp after simulate_fluid_step"
    assert torch.equal(_bits(T.data), _bits(vp0)), "tmp after simulate_fluid_step"
    assert torch.equal(_bits(I.data), _bits(img0)), "dye after advect_color_step"
    del V, T, I, IT, vp, img, vp0, img0
    torch.cuda.empty_cache()
