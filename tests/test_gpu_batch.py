"""Batched scenes (pfs_batch_*): every scene of a batch must end bit-identical
to the reference's CPU solver stepping it alone, whatever else is in the
batch, however often membership changes, with or without graph replay."""
from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from probabilistic_fluid_simulation_amd import scenes  # noqa: E402
from probabilistic_fluid_simulation_amd.batch import Batch  # noqa: E402

from gpu_util import assert_bitexact, oracle_run, to_dev  # noqa: E402

pytestmark = pytest.mark.gpu

MIX = [  # field w, h, dye w, h, K, channels
    (64, 64, 64, 64, 8, 4),       # one tile, dye fused
    (130, 67, 130, 67, 8, 4),     # partial tiles, dye fused
    (200, 150, 97, 301, 8, 4),    # dye of another size: separate dye kernel
    (17, 13, 23, 19, 4, 3),       # tiny, 3-channel dye
    (96, 80, 96, 80, 9, 4),       # K > 8: two chunks
    (256, 128, 256, 128, 4, 4),
    (5, 1, 3, 7, 3, 4),
]


def _scene(i, spec):
    fw, fh, iw, ih, k, ch = spec
    return scenes.random_scene(fw, fh, iw, ih, seed=300 + i, dt=0.6 + 0.05 * i,
                               visc=0.1 + 0.03 * i, iters=k, channels=ch)


def _check(b, slot, scene, orc, steps):
    b.sync()
    got = [t.cpu().numpy() for t in b.buffers(slot)]
    assert_bitexact(got, oracle_run(orc, scene, steps))
    assert b.steps(slot) == steps


@pytest.mark.parametrize("tile", [0, 16, 32, 64], ids=["auto", "t16", "t32", "t64"])
@pytest.mark.parametrize("graphs", [True, False], ids=["graph", "launches"])
def test_batch_mixed_scenes(orc, graphs, tile):
    b = Batch(graphs=graphs)
    sc = [_scene(i, spec) for i, spec in enumerate(MIX)]
    slots = [b.add(*to_dev(s), s.dt, s.visc, s.iters) for s in sc]
    assert b.count == len(MIX)
    b.set_tile(tile)
    assert b.tile == (tile or 16)  # these scenes hold fewer than 128 tiles of 32
    b.step(11)  # with graphs: a captured run of 8 steps, one of 2, one step launched
    for slot, s in zip(slots, sc):
        _check(b, slot, s, orc, 11)
    b.close()


def test_batch_autotune_restores_the_scenes(orc):
    """autotune times every tile side on the live scenes, puts them back as
    they were (buffers and step counts), and forces the fastest side."""
    b = Batch()
    sc = [_scene(i, spec) for i, spec in enumerate(MIX)]
    slots = [b.add(*to_dev(s), s.dt, s.visc, s.iters) for s in sc]
    b.step(3)
    b.sync()
    before = [[t.cpu().numpy().tobytes() for t in b.buffers(sl)] for sl in slots]
    side, ms = b.autotune(reps=3)
    assert side in (16, 32, 64) and ms[side] == min(ms.values()) and min(ms.values()) > 0
    assert b.tile == side
    assert [[t.cpu().numpy().tobytes() for t in b.buffers(sl)] for sl in slots] == before
    assert all(b.steps(sl) == 3 for sl in slots)
    b.step(2)
    for slot, s in zip(slots, sc):
        _check(b, slot, s, orc, 5)
    b.set_tile(0)
    assert b.tile == 16
    b.close()


def test_batch_membership_changes(orc):
    """Scenes join and leave between steps (continuous batching)."""
    b = Batch()
    sc = [_scene(i, spec) for i, spec in enumerate(MIX[:4])]
    slots = {0: b.add(*to_dev(sc[0]), sc[0].dt, sc[0].visc, sc[0].iters),
             1: b.add(*to_dev(sc[1]), sc[1].dt, sc[1].visc, sc[1].iters)}
    taken = {0: 0, 1: 0}
    b.step(3)
    taken[0] += 3
    taken[1] += 3
    slots[2] = b.add(*to_dev(sc[2]), sc[2].dt, sc[2].visc, sc[2].iters)
    taken[2] = 0
    b.step(2)
    for k in taken:
        taken[k] += 2
    _check(b, slots[0], sc[0], orc, taken[0])
    b.remove(slots[0])
    slots[3] = b.add(*to_dev(sc[3]), sc[3].dt, sc[3].visc, sc[3].iters)
    taken[3] = 0
    b.step(4)
    for k in (1, 2, 3):
        taken[k] += 4
    for k in (1, 2, 3):
        _check(b, slots[k], sc[k], orc, taken[k])
    assert b.count == 3
    b.close()


def test_alone_equals_in_the_mix():
    """The served form's invariance: a job's result does not depend on what
    else shares its launches."""
    target = _scene(0, (150, 90, 150, 90, 4, 4))
    alone = Batch()
    s_alone = alone.add(*to_dev(target), target.dt, target.visc, target.iters)
    alone.step(6)
    alone.sync()
    want = [t.cpu().numpy() for t in alone.buffers(s_alone)]
    mix = Batch()
    others = [_scene(i + 1, spec) for i, spec in enumerate(MIX)]
    for o in others[:3]:
        mix.add(*to_dev(o), o.dt, o.visc, o.iters)
    s_mix = mix.add(*to_dev(target), target.dt, target.visc, target.iters)
    for o in others[3:]:
        mix.add(*to_dev(o), o.dt, o.visc, o.iters)
    mix.step(6)
    mix.sync()
    got = [t.cpu().numpy() for t in mix.buffers(s_mix)]
    assert_bitexact(got, want)
    alone.close()
    mix.close()


def test_batch_workspace_and_errors():
    b = Batch()
    s = _scene(0, (64, 64, 64, 64, 8, 4))
    vp, tmp, im, itmp = to_dev(s)
    b.add(vp, tmp, im, itmp, s.dt, s.visc, 8)
    assert b.workspace_bytes == 4 * 4 * 64 * 64  # pw, b and the float2 E plane
    with pytest.raises(Exception):
        b.add(vp, vp, im, itmp, s.dt, s.visc, 8)  # aliasing
    with pytest.raises(Exception):
        b.add(vp, tmp, im, itmp, s.dt, s.visc, 0)  # K = 0 is not batched
    with pytest.raises(Exception):
        b.remove(17)
    b.close()
    np.testing.assert_equal(vp.data.shape, (64, 64, 4))


@pytest.mark.parametrize("substeps", [4, 16])
def test_server_continuous_batching_matches_alone(orc, substeps):
    """Jobs arriving over time, admitted into a running batch and retired as
    they finish, give the same dye as each job run alone; with runs of up to
    16 steps between admission points the batch replays its 8-step graphs
    (dye passes beside the next steps' diffusion), and every job's dye also
    equals the oracle's."""
    from probabilistic_fluid_simulation_amd import serve

    specs = scenes.served_specs(12, rate=400.0, seed=7)
    for sp in specs:  # keep the test small: sides 16..96, 6..40 steps
        side = 16 + sp["fw"] % 81
        sp.update(fw=side, fh=side, iw=side, ih=side, steps=6 + sp["steps"] % 35)
    specs[3].update(iw=specs[3]["fw"] + 5)  # one dye of another size: separate dye kernel
    jobs = serve.jobs_from_specs(specs, iters=4)
    srv = serve.Server(max_substeps=substeps, keep_results=True)
    out = srv.run(jobs)
    srv.close()
    assert out["completed"] == len(jobs)
    assert out["p99_latency_s"] >= out["p50_latency_s"] > 0
    for j in jobs:
        alone = Batch()
        bufs = serve.device_scene(j)
        if substeps > 4:
            host = [t.data.cpu().numpy().copy() for t in bufs]
        slot = alone.add(*bufs, j.dt, j.visc, j.iters)
        alone.step(j.steps)
        alone.sync()
        assert torch.equal(alone.image(slot), j.result), j.job
        alone.close()
        if substeps > 4:
            want = orc.run(*host, j.dt, j.visc, j.iters, j.steps)
            assert j.result.cpu().numpy().tobytes() == want[2].tobytes(), j.job


def test_server_admission_respects_budget():
    from probabilistic_fluid_simulation_amd import serve

    specs = scenes.served_specs(6, rate=1000.0, seed=3)
    for sp in specs:
        sp.update(fw=64, fh=64, iw=64, ih=64, steps=8)
    jobs = serve.jobs_from_specs(specs)
    one = jobs[0].nbytes
    srv = serve.Server(hbm_budget=2 * one, max_substeps=2)
    peak = []
    orig = srv._admit

    def admit(job, now):
        orig(job, now)
        peak.append(len(srv.active))

    srv._admit = admit
    out = srv.run(jobs)
    srv.close()
    assert out["completed"] == 6 and max(peak) <= 2


@pytest.mark.parametrize("iters", [4, 12])
def test_job_bytes_match_what_a_batch_holds(iters):
    """The server's admission count for a job equals its four buffers plus the
    workspace the batch allocates for it."""
    from probabilistic_fluid_simulation_amd import serve

    sp = scenes.served_specs(1, rate=1.0, seed=5)[0]
    sp.update(fw=96, fh=80, iw=120, ih=70, steps=2)
    job = serve.jobs_from_specs([sp], iters=iters)[0]
    bufs = serve.device_scene(job)
    b = Batch()
    b.add(*bufs, job.dt, job.visc, job.iters)
    held = sum(t.data.numel() * 4 for t in bufs) + b.workspace_bytes
    b.close()
    assert job.nbytes == held
