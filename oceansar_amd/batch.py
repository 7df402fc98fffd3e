"""Batching form: many independent scenes of uneven size on one GPU.

The plain path runs one scene after another: every chunk of a small scene
leaves most of the 256 CUs idle and costs a handful of host-side launches.
Here the scenes are spread over S streams (one library context per stream,
since a context owns its synthesis workspace), assigned longest-first to the
least loaded stream, and their chunks are enqueued round-robin so that small
scenes' kernels run beside large ones. The whole schedule can be captured
once into a HIP graph (torch.cuda.CUDAGraph, streams forked and joined inside
the capture) and replayed as one launch. Launch shapes and chunk lengths come
from a timing autotuner with a JSON cache.

Scene semantics are exactly DeviceScene.run's: a batched scene's raw cube
equals the plain run's to float32 rounding of the accumulation order.
"""
import json
import os
import time

import numpy as np

from . import _lib
from .engine import DeviceScene, LaunchShape

SIDES = (128, 256, 512, 1024)
BASELINES = (0., 5., 10., 20.)


def mixed_workload(n=64, seed=4, sides=SIDES, pulses=(512, 4096), max_ch=4, K=16):
    """cfg_text keyword sets of n independent scenes (configs.cfg_text): side
    lengths drawn from `sides` facets at 1 m, U10 uniform in 3-15 m/s, wind
    direction uniform in 0-360 deg, 1..max_ch channels with the along-track
    baselines a prefix of 0/5/10/20 m, pulses uniform in [pulses[0], pulses[1]]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        side = int(rng.choice(np.asarray(sides)))
        nch = int(rng.integers(1, max_ch + 1))
        out.append(dict(L=float(side), dx=1., n_pulses=int(rng.integers(pulses[0], pulses[1] + 1)),
                        K=int(K), U=float(rng.uniform(3., 15.)),
                        wind_dir=float(rng.uniform(0., 360.)), num_ch=nch,
                        b_ati=list(BASELINES[:nch]), seed=int(seed) * 1000 + i))
    return out


def contributions(sc):
    return sc.contributions


class Autotuner:
    """Launch shape (pulses per workgroup P, rows R, window W) and chunk length
    (snapshot intervals per launch group) per scene class, chosen by timing a
    few chunks of each candidate and cached in a JSON file keyed by class.

    The class is (Nx, Ny, channels, polarisations, specular, XTI, K); the cache
    path is `path`, else $OSAR_TUNE_CACHE, else ~/.cache/oceansar_amd/tune.json.
    """

    CHUNKS = (16, 32)

    def __init__(self, path=None):
        self.path = path or os.environ.get('OSAR_TUNE_CACHE') or os.path.join(
            os.path.expanduser('~'), '.cache', 'oceansar_amd', 'tune.json')
        self.cache = {}
        try:
            with open(self.path) as f:
                self.cache = json.load(f)
        except (OSError, ValueError):
            self.cache = {}

    @staticmethod
    def key(sc):
        return '%dx%d_ch%d_%s_spec%d_xti%d_K%d' % (
            sc.Nx, sc.Ny, sc.num_ch, sc.pol, int(sc.spec_enable),
            int(any(b != 0 for b in np.atleast_1d(sc.b_xti))), sc.K)

    @staticmethod
    def candidates(ds, lds_budget=160 * 1024):
        """The default shape, and P in (8, 4, 2) over column chunks of the full
        width down to 1/4 (chunks need a monotonic slant-range table)."""
        sc = ds.sc
        out = [ds.default_shape(lds_budget)]
        for P in (8, 4, 2):
            R = min(8 * (64 // P), sc.Ny)
            for cols in (0, sc.Nx // 2, sc.Nx // 4):
                if cols and (cols < 64 or not ds.sr_monotonic):
                    continue
                W = ds.window_bins(R, cols)
                if ds.lds_bytes(P, sc.num_ch, ds.npol, W) <= lds_budget:
                    s = LaunchShape(P, R, W, cols)
                    if all(s.as_list() != o.as_list() for o in out):
                        out.append(s)
        return out

    def lookup(self, sc):
        e = self.cache.get(self.key(sc))
        if e is None:
            return None
        return LaunchShape(*e['shape']), int(e['chunk'])

    def tune(self, ds, reps=2):
        """Time every candidate on the first chunks of `ds` (all stages) and keep
        the fastest per pulse. Returns (LaunchShape, chunk)."""
        import torch
        hit = self.lookup(ds.sc)
        if hit is not None:
            return hit
        sc = ds.sc
        best = None
        for shape in self.candidates(ds):
            for chunk in self.CHUNKS:
                p1 = min(sc.az_steps, 2 * chunk * sc.K)
                st = ds.start(0, p1, chunk, shape=shape)
                ds.step(st)                          # warm-up (workspace, first launch)
                torch.cuda.synchronize()
                t = time.perf_counter()
                for _ in range(reps):
                    st = ds.start(0, p1, chunk, raw=st.raw, shape=shape)
                    while ds.step(st):
                        pass
                torch.cuda.synchronize()
                per_pulse = (time.perf_counter() - t) / (reps * p1)
                if best is None or per_pulse < best[0]:
                    best = (per_pulse, shape, chunk)
        _, shape, chunk = best
        self.cache[self.key(sc)] = {'shape': shape.as_list(), 'chunk': chunk,
                                    'us_per_pulse': best[0] * 1e6}
        self.save()
        return shape, chunk

    def save(self):
        try:
            os.makedirs(os.path.dirname(self.path), exist_ok=True)
            tmp = self.path + '.tmp'
            with open(tmp, 'w') as f:
                json.dump(self.cache, f, indent=1, sort_keys=True)
            os.replace(tmp, self.path)
        except OSError:
            pass


class BatchRunner:
    """Runs a list of SarScenes on one GPU: sequentially (the plain path),
    concurrently over streams, or as one captured HIP graph."""

    def __init__(self, scenes, n_streams=8, device=0, tuner=None, chunk=None):
        import torch
        self.torch = torch
        self.scenes = list(scenes)
        self.device = device
        self.n_streams = max(1, min(int(n_streams), len(self.scenes)))
        self.ctxs = [_lib.Context(device) for _ in range(self.n_streams)]
        self.streams = [torch.cuda.Stream(device=device) for _ in range(self.n_streams)]
        # longest first onto the least loaded stream
        order = sorted(range(len(self.scenes)), key=lambda i: -contributions(self.scenes[i]))
        load = [0.] * self.n_streams
        self.slot = [0] * len(self.scenes)
        for i in order:
            s = int(np.argmin(load))
            self.slot[i] = s
            load[s] += contributions(self.scenes[i])
        self.ds = []
        self.chunk = []
        for i, sc in enumerate(self.scenes):
            ds = DeviceScene(sc, device=device, ctx=self.ctxs[self.slot[i]])
            c = chunk or 32
            if tuner is not None:
                shape, c = tuner.tune(ds)
                ds.shape = shape
                c = chunk or c
            self.ds.append(ds)
            self.chunk.append(c)
        self.states = None
        self.graph = None

    @property
    def contributions(self):
        return sum(contributions(sc) for sc in self.scenes)

    def _starts(self):
        if self.states is None:
            self.states = [ds.start(0, None, c) for ds, c in zip(self.ds, self.chunk)]
        return self.states

    def _rewind(self, st):
        st.p = st.p0
        st.held = -1
        st.snap_held = -1

    def _zero(self, ds, st):
        ds.zero_c32(st.raw)
        ds.zero_f64(st.rcs_sum)
        if st.spec_energy is not None:
            ds.zero_f64(st.spec_energy)

    def _enqueue_all(self):
        """Round-robin chunk submission: each stream runs its scenes in order;
        every pass over the streams enqueues one chunk per stream."""
        torch = self.torch
        states = self._starts()
        queues = [[i for i in range(len(self.ds)) if self.slot[i] == s]
                  for s in range(self.n_streams)]
        for i, st in enumerate(states):
            self._rewind(st)
        for s in range(self.n_streams):
            with torch.cuda.stream(self.streams[s]):
                for i in queues[s]:
                    self._zero(self.ds[i], states[i])
        pos = [0] * self.n_streams
        active = True
        while active:
            active = False
            for s in range(self.n_streams):
                q = queues[s]
                while pos[s] < len(q):
                    i = q[pos[s]]
                    with torch.cuda.stream(self.streams[s]):
                        more = self.ds[i].step(states[i])
                    if more:
                        active = True
                        break
                    pos[s] += 1

    def run_concurrent(self):
        """Eager run over the streams; returns the makespan in seconds."""
        torch = self.torch
        self._starts()
        torch.cuda.synchronize()
        cur = torch.cuda.current_stream()
        t = time.perf_counter()
        for s in self.streams:
            s.wait_stream(cur)
        self._enqueue_all()
        for s in self.streams:
            cur.wait_stream(s)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    def capture(self):
        """Capture the concurrent schedule (zeroing included) into one graph."""
        torch = self.torch
        self._starts()
        self.run_concurrent()            # sizes every workspace, sets kernel attributes
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        cap = torch.cuda.Stream(device=self.device)
        cap.wait_stream(torch.cuda.current_stream())
        with torch.cuda.graph(g, stream=cap):
            cur = torch.cuda.current_stream()
            for s in self.streams:
                s.wait_stream(cur)
            self._enqueue_all()
            for s in self.streams:
                cur.wait_stream(s)
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self):
        """One graph launch of every scene; returns the makespan in seconds."""
        torch = self.torch
        if self.graph is None:
            self.capture()
        torch.cuda.synchronize()
        t = time.perf_counter()
        self.graph.replay()
        torch.cuda.synchronize()
        return time.perf_counter() - t

    def run_sequential(self, per_scene=None):
        """The plain path: one scene after another on one stream with the
        device's shared context; returns the makespan in seconds. With a list
        `per_scene`, the device is synchronised after every scene and each
        scene's time is appended to it (a separate, slightly slower run)."""
        torch = self.torch
        states = self._starts()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for ds, st in zip(self.ds, states):
            t_scene = time.perf_counter()
            self._rewind(st)
            self._zero(ds, st)
            ctx = ds.ctx
            ds.ctx = _lib.context(self.device)
            try:
                while ds.step(st):
                    pass
            finally:
                ds.ctx = ctx
            if per_scene is not None:
                torch.cuda.synchronize()
                per_scene.append(time.perf_counter() - t_scene)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    def results(self):
        """(raw device tensor, NRCS_avg) of every scene from the last run."""
        return [ds.finish(st) for ds, st in zip(self.ds, self.states)]


def footprint_bytes(sc, chunk=32):
    """Device bytes of one scene run: the resident tables of DeviceScene, the
    chunk buffers of DeviceScene.start and the raw cube. The synthesis
    workspace belongs to the stream's context (workspace_bytes)."""
    from .engine import scene_fields
    n2 = sc.Nx * sc.Ny
    npol = len(sc.pols)
    names = scene_fields(sc)
    tables = 8 * (5 * sc.Nx + sc.Ny) + 8 * n2 + 4 * (sc.Nx + sc.Ny)      # columns, rows, coefs, k
    tables += 8 * (sc.az_steps // sc.K + 2)                            # snapshot times
    if sc.bragg_enable:
        tables += 2 * 4 * n2                                           # scatterer offsets
        tables += 8 * (sc.scat_m.N + sc.scat_p.N + 2 * sc.az_steps)    # extended series
        tables += 4 * sc.az_steps + 8 * len(sc.rcs_refresh)            # schedules
        tables += 16 * len(sc.bragg.spec_k)                            # spectrum table
    bufs = (chunk + 1) * n2 * 4 * len(names)
    if sc.bragg_enable:
        bufs += (chunk + 2) * npol * 2 * n2 * 4
    raw = npol * sc.num_ch * sc.az_steps * sc.rg_samp * 8
    return int(tables + bufs + raw)


def fit_chunk(sc, chunk, free_bytes, nbuf=1, frac=0.9):
    """Largest launch group (snapshot intervals, halving from `chunk`, at least 8)
    whose buffers (nbuf sets of chunk buffers, the resident tables, the raw cube
    and the synthesis workspace) take at most `frac` of `free_bytes`."""
    chunk = int(chunk)
    while chunk > 8 and (nbuf * footprint_bytes(sc, chunk) + workspace_bytes(sc, chunk)
                         > frac * free_bytes):
        chunk //= 2
    return chunk


def workspace_bytes(sc, chunk=32):
    """Synthesis workspace of a context serving scenes of this size: per
    snapshot of a chunk one complex128 plane per packed spectrum (and, for
    Bragg-only scenes, one complex64 plane of the two slopes), and the float64
    dispersion and 1/|k| tables (osar_surface_workspace_bytes)."""
    from .engine import scene_fields, spectra_bytes
    n2 = sc.Nx * sc.Ny
    per = spectra_bytes(scene_fields(sc), sc.surface.choppy,
                        sc.bragg_enable and not sc.spec_enable)
    return int(per * (chunk + 1) * n2 + 16 * n2)


class PersistentBatch:
    """The batching form as persistent launches (osar_echo_accumulate_batch).

    The scenes advance in rounds of one chunk each: the synthesis and Bragg
    refreshes of every scene's next chunk (spread over `n_streams` streams,
    one library context each), then ONE persistent accumulator launch over the
    chunks of all scenes, whose workgroups drain per-XCD queues of work items
    largest descriptor first. Chunk buffers are double-buffered, so a round's
    synthesis and Bragg run beside the previous round's accumulation. All
    scenes share P = 8 pulses per workgroup where every scene's shape fits two
    workgroups per CU (else 4); each item runs the accumulator variant of its
    scene's channel count. `capture()` records
    every round into one HIP graph. A scene's raw cube equals its plain run's
    up to the float32 order of the raw-cube additions.
    """

    def __init__(self, scenes, n_streams=4, device=0, chunk=32, workgroups=0):
        import torch
        self.torch = torch
        self.scenes = list(scenes)
        self.workgroups = int(workgroups)
        for sc in self.scenes:
            if sc.spec_enable or np.any(np.atleast_1d(sc.b_xti) != 0) or sc.num_ch > 4:
                raise ValueError('persistent batches take Bragg-only scenes without '
                                 'cross-track baselines, 1 to 4 channels')
        if len({sc.pol for sc in self.scenes}) != 1:
            raise ValueError('persistent batches take one polarisation mode')
        self.device = device
        self.n_streams = max(1, min(int(n_streams), len(self.scenes)))
        self.ctxs = [_lib.Context(device) for _ in range(self.n_streams)]
        self.streams = [torch.cuda.Stream(device=device) for _ in range(self.n_streams)]
        self.ctx = _lib.context(device)
        order = sorted(range(len(self.scenes)), key=lambda i: -contributions(self.scenes[i]))
        self.slot = [0] * len(self.scenes)
        load = [0.] * self.n_streams
        for i in order:
            s = int(np.argmin(load))
            self.slot[i] = s
            load[s] += contributions(self.scenes[i])
        self.ds = [DeviceScene(sc, device=device, ctx=self.ctxs[self.slot[i]])
                   for i, sc in enumerate(self.scenes)]
        shapes = None
        for P in (8, 4):
            shapes = [ds.default_shape(pulses=(P,)) for ds in self.ds]
            if all(s.pulses_per_block == P for s in shapes):
                break
        if len({s.pulses_per_block for s in shapes}) != 1:
            raise ValueError('no common pulses_per_block (4 or 8) for the batch')
        for ds, s in zip(self.ds, shapes):
            ds.shape = s
        self.chunk = int(chunk)
        # two chunk buffer sets per scene: round r + 1's synthesis and Bragg
        # (buffer set (r + 1) mod 2) run beside round r's accumulation
        self.states = [ds.start(0, None, self.chunk, nbuf=2) for ds in self.ds]
        self.plans = []           # (pinned host, device) plan buffers per round
        self.graph = None

    @property
    def contributions(self):
        return sum(contributions(sc) for sc in self.scenes)

    def _plan(self, r, n):
        torch = self.torch
        nb = int(self.ctx.lib.osar_echo_batch_plan_bytes(n))
        if nb < 0:
            raise ValueError('batch of %d descriptors out of range' % n)
        while len(self.plans) <= r:
            self.plans.append(None)
        have = self.plans[r]
        if have is None or have[0].numel() < nb:
            host = torch.empty(nb, dtype=torch.uint8, pin_memory=True)
            dev = torch.empty(nb, dtype=torch.uint8, device=torch.device('cuda', self.device))
            self.plans[r] = (host, dev)
        return self.plans[r]

    def _enqueue(self):
        """Every round on the streams (synthesis, Bragg) and the current stream
        (one batched accumulation per round)."""
        import ctypes
        torch = self.torch
        cur = torch.cuda.current_stream(self.device)
        for ds, st in zip(self.ds, self.states):
            st.p = st.p0
            st.held = -1
            st.snap_held = -1
            ds.zero_c32(st.raw)
            ds.zero_f64(st.rcs_sum)
        for st in self.states:
            st.k = 0
        for s in self.streams:
            s.wait_stream(cur)          # the zeroing above
        done = [None, None]             # accumulation that last read buffer set r mod 2
        r = 0
        while True:
            if done[r % 2] is not None:
                for s in self.streams:
                    s.wait_event(done[r % 2])
            descs = []
            for i, (ds, st) in enumerate(zip(self.ds, self.states)):
                with torch.cuda.stream(self.streams[self.slot[i]]):
                    d = ds.prepare(st)
                if d is not None:
                    descs.append(d)
            if not descs:
                break
            for s in self.streams:
                cur.wait_stream(s)
            # plan order = queue order: the largest launches first
            descs.sort(key=lambda d: -(d.npulses * d.nx * d.ny))
            arr = (_lib.EchoDesc * len(descs))(*descs)
            host, dev = self._plan(r, len(descs))
            b = _lib.EchoBatchDesc()
            b.n, b.workgroups = len(descs), self.workgroups
            b.descs = ctypes.cast(arr, ctypes.c_void_p)
            b.plan_host, b.plan_dev = host.data_ptr(), dev.data_ptr()
            _lib.check(self.ctx.lib.osar_echo_accumulate_batch(self.ctx.h, ctypes.byref(b),
                                                               _lib.stream_ptr()))
            ev = torch.cuda.Event()
            ev.record(cur)
            done[r % 2] = ev
            r += 1
        for s in self.streams:
            cur.wait_stream(s)
        self.rounds = r

    def run(self):
        """Eager run of every round; returns the makespan in seconds."""
        torch = self.torch
        torch.cuda.synchronize()
        t = time.perf_counter()
        self._enqueue()
        torch.cuda.synchronize()
        return time.perf_counter() - t

    def capture(self):
        """Capture every round into one graph (after an eager run that sizes the
        workspaces and sets the kernels' LDS limits)."""
        torch = self.torch
        self.run()
        g = torch.cuda.CUDAGraph()
        cap = torch.cuda.Stream(device=self.device)
        cap.wait_stream(torch.cuda.current_stream())
        with torch.cuda.graph(g, stream=cap):
            self._enqueue()
        torch.cuda.synchronize()
        self.graph = g
        return g

    def replay(self):
        torch = self.torch
        if self.graph is None:
            self.capture()
        torch.cuda.synchronize()
        t = time.perf_counter()
        self.graph.replay()
        torch.cuda.synchronize()
        return time.perf_counter() - t

    def results(self):
        return [ds.finish(st) for ds, st in zip(self.ds, self.states)]
