"""Serving simulation jobs with continuous batching.

The reference runs one simulation per process (src/main.cpp).  Here a
request loop admits jobs as they arrive, keeps every admitted job's state
resident on the device, and advances all of them together through a
``Batch`` (pfs_batch_*): runs of batched steps are queued ahead of the
device, jobs that finish in a run leave the batch behind it (their dye
copied out on the same stream), and waiting jobs are admitted as slots and
the HBM budget allow.  Each job's result is bit-identical to running it
alone: scenes never interact inside the batched kernels.

Job inputs follow the 16K-cell workload's recipe (Taylor-Green cells and a
disk of dye), generated on the device at admission so that the request
path does not wait on host numpy.
"""
from __future__ import annotations

import collections
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
    rejected: bool = False  # larger than the server's whole HBM budget

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


class _Device:
    """The GPU side of a Server: the batch's stream, a side stream on which
    admitted jobs' inputs are generated while the batch runs, and events."""

    def __init__(self, batch: Batch):
        self.stream = torch.cuda.ExternalStream(batch.stream_handle)
        self.side = torch.cuda.Stream()

    def make(self, job: "Job"):
        """The job's buffers, generated on the side stream; the batch's
        stream waits for them before its next step."""
        with torch.cuda.stream(self.side):
            bufs = device_scene(job)
        ready = torch.cuda.Event()
        ready.record(self.side)
        self.stream.wait_event(ready)
        return bufs

    def mark(self):
        """An event after everything queued so far on the batch's stream."""
        ev = torch.cuda.Event()
        ev.record(self.stream)
        return ev

    def copy_out(self, t: torch.Tensor) -> torch.Tensor:
        """A copy of t made on the batch's stream after the work queued there."""
        with torch.cuda.stream(self.stream):
            return t.clone()


class Server:
    """Admission + continuous batching over one Batch.

    hbm_budget: bytes the resident jobs may use (a job larger than the whole
    budget is rejected, never queued); max_active: slot limit (>= 1);
    max_substeps: longest run of steps between admission points; depth:
    runs of steps queued on the device ahead of the host (>= 1), so that the
    device never waits for the host to retire and admit jobs.

    Nothing synchronises the device with the host per step: a run of n
    steps is queued, jobs that finish in it are removed from later steps and
    their results copied on the batch's stream right behind it, and an event
    marks its end.  The host then queues the next run (up to `depth` ahead)
    and meanwhile admits arrivals, whose inputs are generated on a side
    stream.  A job counts as finished, and its memory as free, when the host
    sees its run's event complete.

    `batch` and `device` replace the Batch and its GPU side (tests drive the
    admission logic with stand-ins on the CPU)."""

    def __init__(self, hbm_budget: float = 200e9, max_active: int = 256, max_substeps: int = 16,
                 keep_results: bool = False, depth: int = 2, batch=None, device=None):
        if not hbm_budget > 0:
            raise ValueError("hbm_budget must be positive")
        if max_active < 1 or max_substeps < 1 or depth < 1:
            raise ValueError("max_active, max_substeps and depth must be >= 1")
        self.batch = batch if batch is not None else Batch()
        self.dev = device if device is not None else _Device(self.batch)
        self.hbm_budget = hbm_budget
        self.max_active = max_active
        self.max_substeps = max_substeps
        self.keep_results = keep_results
        self.depth = depth
        self.active: Dict[int, tuple] = {}  # slot -> (job, steps not yet queued, buffers)
        self.used = 0  # bytes of admitted jobs whose last run has not been seen to end
        self.peak_active = 0

    def _admit(self, job: Job, now: float) -> None:
        bufs = self.dev.make(job)
        slot = self.batch.add(*bufs, job.dt, job.visc, job.iters, sync=False)
        job.admitted = now
        self.active[slot] = (job, job.steps, bufs)
        self.used += job.nbytes
        self.peak_active = max(self.peak_active, len(self.active))

    def _queue_run(self) -> tuple:
        """Queues the steps up to the next completion (at most max_substeps)
        for every resident job; returns (end event, finishing jobs)."""
        n = min(min(rem for _, rem, _ in self.active.values()), self.max_substeps)
        self.batch.step(n)
        done = []
        for slot in list(self.active):
            job, rem, bufs = self.active[slot]
            if rem == n:
                res = self.dev.copy_out(self.batch.image(slot)) if self.keep_results else None
                bufs = self.batch.remove(slot, hold=False)
                del self.active[slot]
                done.append((job, bufs, res))
            else:
                self.active[slot] = (job, rem - n, bufs)
        return self.dev.mark(), done

    def run(self, jobs: List[Job], clock=time.perf_counter, sleep=time.sleep) -> dict:
        for j in jobs:
            j.rejected = j.nbytes > self.hbm_budget
        pending = collections.deque(sorted((j for j in jobs if not j.rejected),
                                           key=lambda j: j.arrival))
        inflight: collections.deque = collections.deque()
        t0 = clock()
        while pending or self.active or inflight:
            now = clock() - t0
            while (pending and pending[0].arrival <= now
                   and len(self.active) < self.max_active
                   and self.used + pending[0].nbytes <= self.hbm_budget):
                self._admit(pending.popleft(), now)
            if self.active and len(inflight) < self.depth:
                inflight.append(self._queue_run())
                continue
            if inflight:
                ev, done = inflight[0]
                if not ev.query():
                    nxt = pending[0] if pending else None
                    if (nxt is not None and len(self.active) < self.max_active
                            and self.used + nxt.nbytes <= self.hbm_budget):
                        # The next job fits: admit it on arrival, while the
                        # device works.
                        sleep(min(max(0.0, nxt.arrival - (clock() - t0)), 1e-4))
                    else:
                        ev.synchronize()  # only the run's end can free room
                    continue
                inflight.popleft()
                now = clock() - t0
                for job, _bufs, res in done:
                    job.finished = now
                    job.result = res
                    self.used -= job.nbytes
                continue
            # Idle: nothing resident, nothing queued; the next job is due later.
            sleep(max(0.0, pending[0].arrival - (clock() - t0)))
        self.batch.sync()
        return summarize(jobs, clock() - t0)

    def close(self) -> None:
        self.batch.close()


def summarize(jobs: List[Job], makespan: float) -> dict:
    lat = np.array([j.latency for j in jobs if j.latency is not None])
    cells = sum(j.fw * j.fh * j.steps for j in jobs)
    return {"jobs": len(jobs), "completed": int(lat.size),
            "rejected": sum(1 for j in jobs if j.rejected), "makespan_s": makespan,
            "jobs_per_s": lat.size / makespan if makespan > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
            "cell_steps_per_s": cells / makespan if makespan > 0 else 0.0}
