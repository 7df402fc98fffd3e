"""Serving form: simulation jobs of different sizes arriving over time.

A job is a prepared scene (SarScene: cfg, geometry, wave spectrum and
coefficients, scatterer series - the client's host work) with an arrival
time. The server keeps at most S jobs on the device, one per slot; a slot is a
HIP stream with its own library context. Jobs are admitted in arrival order
(skipping over a job that does not fit yet, so smaller ones behind it can use
a free slot) while a slot is free and the device footprint of the admitted
jobs plus the slot workspaces stays under the HBM budget. Admitted jobs
upload their tables, then advance one chunk at a time, round-robin, with at
most `ahead` chunks enqueued per job, so that a job admitted later starts at
once on its own stream and overlaps on the device with the running ones
(continuous batching, not a queue). A finished job's raw cube goes to pinned
host memory with an asynchronous copy; its latency is arrival -> result on
the host. `serve_sequential` runs the same jobs one at a time on the plain
path for comparison.

Each job's result depends only on its scene: the same kernels, launch shape
and chunking alone or in a mix (to float32 rounding of the order of the
atomic accumulation).
"""
import collections
import time

import numpy as np

from . import _lib
from .batch import footprint_bytes, workspace_bytes
from .engine import DeviceScene


class Job:
    """One simulation request: `scene` (SarScene) arriving `arrival` seconds
    after the server starts."""

    def __init__(self, jid, scene, arrival=0.):
        self.id = jid
        self.scene = scene
        self.arrival = float(arrival)
        self.t_admit = None
        self.t_done = None
        self.raw = None          # host complex64 [pol][ch][pulse][rg_samp]
        self.nrcs = None
        # device-side state while active
        self._ds = None
        self._st = None
        self._events = collections.deque()
        self._host = None
        self._copy_ev = None
        self._steps_left = True

    @property
    def latency(self):
        return None if self.t_done is None else self.t_done - self.arrival

    @property
    def contributions(self):
        return self.scene.contributions


def poisson_arrivals(n, rate, seed=5):
    """n arrival times (s) of a Poisson process with `rate` jobs per second."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1. / rate, size=n))


def stats(jobs, t_first=None):
    lat = np.array([j.latency for j in jobs], dtype=np.float64)
    w = float(sum(j.contributions for j in jobs))
    t0 = min(j.arrival for j in jobs) if t_first is None else t_first
    span = max(j.t_done for j in jobs) - t0
    return {'jobs': len(jobs), 'p50_latency_s': float(np.percentile(lat, 50)),
            'p99_latency_s': float(np.percentile(lat, 99)),
            'mean_latency_s': float(lat.mean()), 'makespan_s': float(span),
            'contributions': w, 'contributions_per_s': w / span}


class Server:
    def __init__(self, n_slots=4, hbm_budget=200e9, device=0, chunk=16, ahead=2):
        import torch
        self.torch = torch
        self.device = device
        self.n_slots = int(n_slots)
        self.budget = float(hbm_budget)
        self.chunk = int(chunk)
        self.ahead = max(1, int(ahead))
        self.ctxs = [_lib.Context(device) for _ in range(self.n_slots)]
        self.streams = [torch.cuda.Stream(device=device) for _ in range(self.n_slots)]
        self.ws = [0] * self.n_slots         # workspace bytes each slot's context may hold

    def _need(self, slot, job):
        grow = max(0, workspace_bytes(job.scene, self.chunk) - self.ws[slot])
        return footprint_bytes(job.scene, self.chunk) + grow

    def _admit(self, slot, job, now):
        torch = self.torch
        with torch.cuda.stream(self.streams[slot]):
            ds = DeviceScene(job.scene, device=self.device, ctx=self.ctxs[slot])
            job._ds = ds
            job._st = ds.start(0, None, self.chunk)
        self.ws[slot] = max(self.ws[slot], workspace_bytes(job.scene, self.chunk))
        job.t_admit = now

    def _advance(self, slot, job):
        """Enqueue one more chunk if fewer than `ahead` are in flight."""
        torch = self.torch
        while job._events and job._events[0].query():
            job._events.popleft()
        if not job._steps_left or len(job._events) >= self.ahead:
            return False
        with torch.cuda.stream(self.streams[slot]):
            job._steps_left = job._ds.step(job._st)
            if job._steps_left:
                ev = torch.cuda.Event()
                ev.record()
                job._events.append(ev)
            else:
                # all chunks enqueued: copy the raw cube out asynchronously
                raw = job._st.raw
                job._host = torch.empty(raw.shape, dtype=raw.dtype, pin_memory=True)
                job._host.copy_(raw, non_blocking=True)
                job._copy_ev = torch.cuda.Event()
                job._copy_ev.record()
        return True

    def _finished(self, job):
        return (not job._steps_left) and job._copy_ev is not None and job._copy_ev.query()

    def serve(self, jobs, clock=time.perf_counter):
        """Run `jobs` (arrival times relative to the call); returns stats()."""
        torch = self.torch
        pending = collections.deque(sorted(jobs, key=lambda j: j.arrival))
        queue = []
        active = {}                 # slot -> job
        used = 0
        t0 = clock()
        while pending or queue or active:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            # admission: arrival order, skipping jobs that do not fit yet
            free = [s for s in range(self.n_slots) if s not in active]
            for job in list(queue):
                if not free:
                    break
                slot = free[0]
                need = self._need(slot, job)
                if used + need <= self.budget or not active:
                    used += footprint_bytes(job.scene, self.chunk) + max(
                        0, workspace_bytes(job.scene, self.chunk) - self.ws[slot])
                    self._admit(slot, job, now)
                    active[slot] = job
                    queue.remove(job)
                    free.pop(0)
            progressed = False
            for slot, job in list(active.items()):
                progressed |= self._advance(slot, job)
                if self._finished(job):
                    job.raw = job._host.numpy()
                    _, job.nrcs = job._ds.finish(job._st)
                    job.t_done = clock() - t0
                    used -= footprint_bytes(job.scene, self.chunk)
                    job._ds = job._st = job._host = None
                    job._events.clear()
                    del active[slot]
                    progressed = True
            if not progressed:
                # nothing to enqueue: wait a little for the device or the next arrival
                wait = 2e-4
                if pending and not active:
                    wait = max(0., pending[0].arrival - (clock() - t0))
                time.sleep(min(wait, 5e-3))
        torch.cuda.synchronize()
        return stats(jobs, 0.)


def serve_sequential(jobs, device=0, chunk=16, clock=time.perf_counter):
    """The plain path for the same arrivals: one job at a time, in arrival
    order, on one stream; a job starts when it has arrived and the previous
    one has returned its result."""
    import torch
    t0 = clock()
    for job in sorted(jobs, key=lambda j: j.arrival):
        now = clock() - t0
        if now < job.arrival:
            time.sleep(job.arrival - now)
        job.t_admit = clock() - t0
        ds = DeviceScene(job.scene, device=device)
        raw, job.nrcs = ds.run(chunk_snaps=chunk)
        job.raw = raw.cpu().numpy()
        job.t_done = clock() - t0
    torch.cuda.synchronize()
    return stats(jobs, 0.)
