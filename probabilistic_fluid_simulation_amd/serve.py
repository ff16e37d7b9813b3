"""Serving simulation jobs with continuous batching.

The reference runs one simulation per process (src/main.cpp).  Here a
request loop admits jobs as they arrive, keeps every admitted job's state
resident on the device, and advances all of them together through a
``Batch`` (pfs_batch_*): after every batched substep, finished jobs are
retired (their dye copied out) and waiting jobs are admitted as slots and
the HBM budget allow.  Each job's result is bit-identical to running it
alone: scenes never interact inside the batched kernels.

Job inputs follow the 16K-cell workload's recipe (Taylor-Green cells and a
disk of dye), generated on the device at admission so that the request
path does not wait on host numpy.
"""
from __future__ import annotations

import dataclasses
import math
import time
from typing import Dict, List, Optional

import numpy as np
import torch

from . import scenes
from .batch import Batch
from .fluid import VPField


@dataclasses.dataclass
class Job:
    job: int
    arrival: float  # seconds after the stream starts
    fw: int
    fh: int
    iw: int
    ih: int
    steps: int
    seed: int
    iters: int = 4
    dt: float = 1.0 / 60.0
    visc: float = 0.1
    # filled in by the server
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[torch.Tensor] = None

    @property
    def nbytes(self) -> int:
        """Device bytes while resident: vp, tmp (16 B/cell each), image and
        itmp (16 B/pixel each) and the batch's scene workspace (16 B/cell for
        K <= 8, 32 B/cell beyond; pfs_batch_workspace_bytes)."""
        ws = 16 if self.iters <= 8 else 32
        return (32 + ws) * self.fw * self.fh + 32 * self.iw * self.ih

    @property
    def latency(self) -> Optional[float]:
        return None if self.finished is None else self.finished - self.arrival


def device_scene(job: Job, device: str = "cuda"):
    """The job's four buffers, made on the device: 4 Taylor-Green cells with
    wavenumber, phases and amplitudes from SplitMix64(seed) as in
    scenes.taylor_green (evaluated in float64 on the device), a disk of dye of
    radius 0.3 with x/y ramps (scenes.dye_disk), vp = (u, v, -1, 1) and
    tmp = (-1, -1, -1, 1) as src/main.cpp:170-195 initialises them."""
    r = scenes.uniform(job.seed, 20, stream=1).reshape(4, 5)
    f64 = dict(dtype=torch.float64, device=device)
    x = (torch.arange(job.fw, **f64) + 0.5) / job.fw
    y = (torch.arange(job.fh, **f64) + 0.5) / job.fh
    Y, X = torch.meshgrid(y, x, indexing="ij")
    u = torch.zeros_like(X)
    v = torch.zeros_like(X)
    for k in range(4):
        kw = 1 + int(r[k, 0] * 4)
        a = 0.1 + 0.9 * r[k, 3]
        sx = 2 * math.pi * kw * X + 2 * math.pi * r[k, 1]
        sy = 2 * math.pi * kw * Y + 2 * math.pi * r[k, 2]
        u += a * torch.sin(sx) * torch.cos(sy)
        v -= a * torch.cos(sx) * torch.sin(sy)
    s = torch.maximum(u.abs().max(), v.abs().max()).clamp_min(1e-30)
    vp = torch.empty((job.fh, job.fw, 4), dtype=torch.float32, device=device)
    vp[..., 0] = (u / s).float()
    vp[..., 1] = (v / s).float()
    vp[..., 2] = -1.0
    vp[..., 3] = 1.0
    tmp = torch.full_like(vp, -1.0)
    tmp[..., 3] = 1.0
    xi = (torch.arange(job.iw, **f64) + 0.5) / job.iw
    yi = (torch.arange(job.ih, **f64) + 0.5) / job.ih
    Yi, Xi = torch.meshgrid(yi, xi, indexing="ij")
    im = torch.empty((job.ih, job.iw, 4), dtype=torch.float32, device=device)
    im[..., 0] = (((Xi - 0.5) ** 2 + (Yi - 0.5) ** 2) <= 0.09).float()
    im[..., 1] = Xi.float()
    im[..., 2] = Yi.float()
    im[..., 3] = 1.0
    itmp = torch.zeros_like(im)
    return tuple(VPField(t) for t in (vp, tmp, im, itmp))


def jobs_from_specs(specs: List[dict], iters: int = 4) -> List[Job]:
    return [Job(job=s["job"], arrival=s["arrival"], fw=s["fw"], fh=s["fh"], iw=s["iw"],
                ih=s["ih"], steps=s["steps"], seed=s["seed"], iters=iters) for s in specs]


class Server:
    """Admission + continuous batching over one Batch.

    hbm_budget: bytes the resident jobs may use; max_active: slot limit;
    max_substeps: longest run of steps between admission checks."""

    def __init__(self, hbm_budget: float = 200e9, max_active: int = 256, max_substeps: int = 16,
                 keep_results: bool = False):
        self.batch = Batch()
        self.hbm_budget = hbm_budget
        self.max_active = max_active
        self.max_substeps = max_substeps
        self.keep_results = keep_results
        self.active: Dict[int, tuple] = {}  # slot -> (job, remaining, buffers)
        self.used = 0

    def _admit(self, job: Job, now: float) -> None:
        bufs = device_scene(job)
        slot = self.batch.add(*bufs, job.dt, job.visc, job.iters)
        job.admitted = now
        self.active[slot] = (job, job.steps, bufs)
        self.used += job.nbytes

    def _retire(self, slot: int, now: float) -> None:
        job, _, bufs = self.active.pop(slot)
        if self.keep_results:
            job.result = self.batch.image(slot).clone()
        self.batch.remove(slot)
        job.finished = now
        self.used -= job.nbytes

    def run(self, jobs: List[Job], clock=time.perf_counter) -> dict:
        pending = sorted(jobs, key=lambda j: j.arrival)
        i = 0
        t0 = clock()
        while i < len(pending) or self.active:
            now = clock() - t0
            while (i < len(pending) and pending[i].arrival <= now
                   and len(self.active) < self.max_active
                   and self.used + pending[i].nbytes <= self.hbm_budget):
                self._admit(pending[i], now)
                i += 1
            if not self.active:
                if i < len(pending):
                    time.sleep(max(0.0, pending[i].arrival - (clock() - t0)))
                continue
            # Advance every resident job to the next completion, at most
            # max_substeps steps, so that arrivals wait at most that long.
            n = min(min(rem for _, rem, _ in self.active.values()), self.max_substeps)
            self.batch.step(n)
            self.batch.sync()
            now = clock() - t0
            for slot in list(self.active):
                job, rem, bufs = self.active[slot]
                if rem - n == 0:
                    self._retire(slot, now)
                else:
                    self.active[slot] = (job, rem - n, bufs)
        self.batch.sync()
        return summarize(jobs, clock() - t0)

    def close(self) -> None:
        self.batch.close()


def summarize(jobs: List[Job], makespan: float) -> dict:
    lat = np.array([j.latency for j in jobs if j.latency is not None])
    cells = sum(j.fw * j.fh * j.steps for j in jobs)
    return {"jobs": len(jobs), "completed": int(lat.size), "makespan_s": makespan,
            "jobs_per_s": lat.size / makespan if makespan > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
            "cell_steps_per_s": cells / makespan if makespan > 0 else 0.0}
