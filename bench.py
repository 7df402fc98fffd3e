"""Throughput of the OCEANSAR raw-data generation hot path on one MI355X.

A step is one full raw-data simulation of the workload: surface synthesis
(FP64 FFTs) every K pulses, Bragg facet RCS at every refresh, and the fused
echo accumulation of every facet into every pulse's range line for every
channel. Inputs (wave coefficients, geometry tables, closure series) are
resident in HBM before the timed region; the raw cube stays on the device
for `value`, and the same steps ending with the raw cube in pinned host
memory give `value_with_host_copy`.

Metric: facet-pulse contributions per second (one facet's echo deposited into
one pulse's range line for one channel and polarisation), and the p50/p99
scene latency of the serving form (Poisson arrivals of mixed-size jobs),
which the default run appends to the same line. The CPU baseline leg times
the CPU restatement on bounded pulse samples and checks the GPU run's raw
data on the same pulses against it (rel. RMS, phase RMS).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload large]
"""
import argparse
import json
import os
import re
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 vector peak (published)
HBM_PEAK_GBS = 8000.0
_T0 = time.perf_counter()


def progress(msg):
    """One line per finished stage on stderr (the JSON line alone goes to stdout)."""
    print('[bench %7.1f s] %s' % (time.perf_counter() - _T0, msg), file=sys.stderr, flush=True)


def flop_per_contribution(n_ch, n_pol):
    """Algorithmic FP32 FLOP of one contribution: 38 per (facet, pulse, channel, pol)
    (8 taps x real*complex MAC = 32, channel phasor 6) plus 35 per (facet, pulse, pol)
    (residual range, binning, the two Bragg terms, surface phasor, beam) shared by
    the channels."""
    return 38. + 35. / n_ch


def configs_names():
    from oceansar_amd.configs import WORKLOADS
    return WORKLOADS


def measured_traffic(sc, shape_repr, pulses_per_launch):
    """HBM bytes per echo launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary of this workload (counters cannot be read from inside the timed run).
    Returns (bytes or None, source file); None when the summary was taken at a
    different launch shape or launch size."""
    prof = os.path.join(ROOT, 'profiles')
    try:
        names = [n for n in os.listdir(prof)
                 if n == 'raw_generation_%dx%d_%dpulses_%dch_%s_hbm_traffic.md'
                 % (sc.Nx, sc.Ny, sc.az_steps, sc.num_ch, sc.pol)]
    except OSError:
        return None, None
    for n in names:
        text = open(os.path.join(prof, n)).read()
        m = re.search(r'traffic_bytes_per_launch echo_accumulate_kernel = ([0-9.eE+-]+)', text)
        s = re.search(r'pulses_per_launch = (\d+), launch_shape = (.*)', text)
        if m and s and int(s.group(1)) == pulses_per_launch and s.group(2).strip() == shape_repr:
            return float(m.group(1)), 'profiles/' + n
    return None, None


def host_cores():
    """(threads this process may use, os.cpu_count(), CPU model): the affinity
    set, capped by a cgroup CPU quota when one is set."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else os.cpu_count()
    try:
        q, per = open('/sys/fs/cgroup/cpu.max').read().split()[:2]
        if q != 'max':
            n = min(n, max(1, int(float(q) / float(per) + 0.5)))
    except (OSError, ValueError):
        pass
    model = ''
    try:
        for ln in open('/proc/cpuinfo'):
            if ln.startswith('model name'):
                model = ln.split(':', 1)[1].strip()
                break
    except OSError:
        pass
    return int(n), os.cpu_count(), model


def cpu_baseline(sc, raw_gpu, sample_pulses, sample_pulses_1t):
    """The CPU restatement of the reference loop on host cores, over bounded
    pulse samples of the same scene: numpy FP64 surface synthesis of the
    sample's snapshots and numpy FP64 Bragg refreshes (the reference's own
    numpy stages), then the C pulse loop (the numba nopython stand-in) at
    every available core and at one thread. The all-core sample is also the
    accuracy check of the GPU run: the bench's raw cube on the same pulses
    against this FP64 result."""
    import oracle.c_oracle as co
    import oracle.sar as osar
    threads, n_os, model = host_cores()
    K = sc.K
    p0 = (sc.az_steps // 2 // K) * K - K // 2          # crosses a snapshot and a refresh
    p1 = p0 + sample_pulses
    n_snap = (p1 - 1) // K - p0 // K + 2
    snaps = osar.SnapshotCache(sc, keep=n_snap + 2)
    tim = {}
    ref, _ = co.simulate_c(sc, p0, p1, snaps=snaps, threads=threads, timings=tim)
    progress('cpu baseline: %d pulses at %d threads done' % (p1 - p0, threads))
    npol = len(sc.pols)
    per_pulse = sc.Nx * sc.Ny * sc.num_ch * npol
    w = per_pulse * (p1 - p0)
    t_all = tim['synth_s'] + tim['bragg_s'] + tim['loop_s']
    # one thread: the C loop only, over a shorter window of the same snapshots
    tim1 = {}
    co.simulate_c(sc, p0, p0 + sample_pulses_1t, snaps=snaps, threads=1, timings=tim1)
    progress('cpu baseline: %d pulses at 1 thread done' % sample_pulses_1t)
    w1 = per_pulse * sample_pulses_1t
    # accuracy of the GPU run on the sample pulses (noise-free, uncalibrated raw)
    got = raw_gpu[:, :, p0:p1].cpu().numpy().astype(np.complex128)
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    mag = np.abs(ref)
    keep = mag > mag.max() * 10 ** (-30. / 20.)
    ph = float(np.sqrt(np.mean(np.angle(got[keep] * np.conj(ref[keep])) ** 2)))
    n_refresh = len(set(int(r) for r in sc.rcs_slot[p0:p1]))
    loop_1t = w1 / tim1['loop_s']
    cpu = {
        'value': w / t_all, 'unit': 'contributions/s', 'cores': threads, 'kind': 'port',
        'sample': 'pulses [%d, %d) of %d: %d snapshots synthesised (numpy FP64), %d Bragg '
                  'refreshes (numpy FP64), the C pulse loop at %d threads; %.1f s'
                  % (p0, p1, sc.az_steps, n_snap, n_refresh, threads, t_all),
        'stage_s': {k: round(v, 3) for k, v in tim.items()},
        'loop_only_value': w / tim['loop_s'],
        'loop_only_value_1_thread': loop_1t,
        'sample_1_thread': 'pulses [%d, %d), C loop only, %.1f s'
                           % (p0, p0 + sample_pulses_1t, tim1['loop_s']),
        'host': {'threads_used': threads, 'os_cpu_count': n_os, 'cpu_model': model},
    }
    acc = {'rel_rms': err, 'phase_rms_rad': ph, 'against': 'C FP64 oracle (oracle/c), pulses '
           '[%d, %d) of the timed run, %d channels' % (p0, p1, sc.num_ch),
           'bound_rel_rms': 1e-4, 'bound_phase_rms_rad': 1e-3}
    return cpu, acc


def _scenes(kws, tag, device_init=None):
    """Scenes of the given cfg keywords; device_init (a device index): their
    spectra and coefficients are built in HBM (osar_surface_init)."""
    from oceansar_amd import configs
    d = tempfile.mkdtemp(prefix=tag)
    t = time.perf_counter()
    scenes = [configs.load_scene(configs.cfg_text(**kw), d, name='s%d.cfg' % i,
                                 device_init=device_init)
              for i, kw in enumerate(kws)]
    return scenes, time.perf_counter() - t


def bench_batch(args):
    """64 independent scenes (sides 128-1024 at 1 m, 512-4096 pulses, 1-4
    channels, U10 3-15 m/s, any wind direction; seed 4) with every input
    resident: the plain path one scene after another, the same scenes
    concurrently over streams, their captured graph replayed as one launch,
    and the persistent form (one persistent accumulator launch per round of
    chunks, every round in one graph). value = contributions / makespan of the
    persistent form's graph replay."""
    import torch
    from oceansar_amd import batch
    n = args.jobs or 64
    scenes, t_host = _scenes(batch.mixed_workload(n=n, seed=4), 'batch')
    tuner = batch.Autotuner()
    t = time.perf_counter()
    br = batch.BatchRunner(scenes, n_streams=args.streams, tuner=tuner)
    t_setup = time.perf_counter() - t
    W = br.contributions
    br.run_sequential()                       # warm-up: workspaces, kernel attributes
    seq = min(br.run_sequential() for _ in range(2))
    # the plain path per side length (synchronised after each scene): where the
    # plain path leaves the device underused, batching can gain
    times = []
    br.run_sequential(per_scene=times)
    per_side = {}
    for sc, dt in zip(scenes, times):
        c = per_side.setdefault(str(sc.Nx), {'scenes': 0, 'time_s': 0., 'contributions': 0})
        c['scenes'] += 1
        c['time_s'] += dt
        c['contributions'] += sc.contributions
    tot = sum(times)
    for c in per_side.values():
        c['time_share'] = c['time_s'] / tot
        c['contributions_per_s'] = c['contributions'] / c['time_s']
    conc = min(br.run_concurrent() for _ in range(2))
    t = time.perf_counter()
    br.capture()
    t_capture = time.perf_counter() - t
    graph = min(br.replay() for _ in range(3))
    torch.cuda.synchronize()
    # persistent form: per round of chunks, one persistent accumulator launch
    # over every scene (osar_echo_accumulate_batch), all rounds in one graph
    pb = batch.PersistentBatch(scenes, n_streams=args.streams, chunk=args.batch_chunk)
    t = time.perf_counter()
    pb.capture()
    t_pcapture = time.perf_counter() - t
    pers = min(pb.replay() for _ in range(3))
    pers_eager = pb.run()
    torch.cuda.synchronize()
    rounds = pb.rounds
    del pb
    torch.cuda.empty_cache()
    return {
        'metric': 'facet-pulse contributions/s (batching form: %d mixed scenes, persistent '
                  'launches in one graph)' % n,
        'value': W / pers, 'unit': 'contributions/s', 'n_gpus': 1, 'higher_is_better': True,
        'dtype': 'f32 (FP64 range/phase/bins)', 'data': 'synthetic (seed 4)',
        'config': {'workload': '%d scenes, sides 128-1024 at 1 m, 512-4096 pulses, 1-4 channels, '
                               'U10 3-15 m/s, K=16' % n,
                   'contributions': W, 'streams': br.n_streams,
                   'sides': sorted(set(sc.Nx for sc in scenes)),
                   'tuned_classes': len(tuner.cache)},
        'makespan_s': {'plain_sequential': seq, 'concurrent_streams': conc, 'graph_replay': graph,
                       'persistent_eager': pers_eager, 'persistent_graph': pers},
        'persistent_rounds': rounds,
        'plain_value': W / seq, 'speedup_vs_plain': seq / pers,
        'stream_graph_speedup_vs_plain': seq / graph,
        'plain_per_side': per_side,
        'host_s': {'scene_setup': t_host, 'upload_and_tune': t_setup, 'graph_capture': t_capture,
                   'persistent_setup_and_capture': t_pcapture},
    }


def bench_serve(args):
    """Serving form: jobs sized as the batching form plus 10 % at 2048^2 with
    2 channels (seed 5), Poisson arrivals at 0.8 x the measured plain capacity,
    admitted by HBM footprint under 200 GB. Reports p50/p99 latency (arrival ->
    result on the host) and throughput, beside the same arrivals run one job at
    a time on the plain path."""
    import numpy as np
    import torch
    from oceansar_amd import batch, serve
    n = args.jobs or 48
    kws = batch.mixed_workload(n=n, seed=5)
    rng = np.random.default_rng(55)
    for kw in kws:
        if rng.uniform() < 0.1:
            kw.update(L=2048., dx=1., num_ch=2, b_ati=[0., 5.])
    # job surfaces are built on the device: their spectra never leave HBM
    scenes, t_host = _scenes(kws, 'serve', device_init=None if args.host_init else 0)
    W = sum(sc.contributions for sc in scenes)
    # plain capacity: the jobs back to back (every job arrived at t = 0)
    warm = [serve.Job(i, sc, 0.) for i, sc in enumerate(scenes[:2])]
    serve.serve_sequential(warm)
    cap_jobs = [serve.Job(i, sc, 0.) for i, sc in enumerate(scenes)]
    cap = serve.serve_sequential(cap_jobs)
    capacity = W / cap['makespan_s']
    rate = 0.8 * capacity / (W / n)                 # jobs per second
    arrivals = serve.poisson_arrivals(n, rate, seed=5)
    srv = serve.Server(n_slots=args.streams, hbm_budget=200e9)
    srv.serve([serve.Job(i, sc, 0.) for i, sc in enumerate(scenes[:2])])   # warm-up
    jobs = [serve.Job(i, sc, a) for i, (sc, a) in enumerate(zip(scenes, arrivals))]
    got = srv.serve(jobs)
    progress('serving: %d jobs served' % n)
    seq_jobs = [serve.Job(i, sc, a) for i, (sc, a) in enumerate(zip(scenes, arrivals))]
    plain = serve.serve_sequential(seq_jobs)
    torch.cuda.synchronize()
    progress('serving: the same arrivals on the plain path done')
    return {
        'metric': 'p99 scene latency (serving form: %d jobs, Poisson arrivals at 0.8 x plain capacity)' % n,
        'value': got['p99_latency_s'], 'unit': 's', 'n_gpus': 1, 'higher_is_better': False,
        'dtype': 'f32 (FP64 range/phase/bins)', 'data': 'synthetic (seed 5)',
        'config': {'workload': '%d jobs: sides 128-1024 at 1 m (10%% 2048^2, 2 ch), 512-4096 pulses, '
                               '1-4 channels; HBM budget 200 GB; %d slots' % (n, srv.n_slots),
                   'contributions': W, 'arrival_rate_per_s': rate},
        'served': got, 'plain_sequential': plain,
        'p50_speedup_vs_plain': plain['p50_latency_s'] / got['p50_latency_s'],
        'p99_speedup_vs_plain': plain['p99_latency_s'] / got['p99_latency_s'],
        'throughput_vs_plain': got['contributions_per_s'] / plain['contributions_per_s'],
        'plain_capacity': capacity, 'host_scene_setup_s': t_host,
        'surface_init': 'host numpy' if args.host_init else 'device (osar_surface_init)',
    }


def bench_insar(args):
    """InSAR L1b processing at the large workload's SLC size: 4 channels x 7555
    lines x 624 range samples (the valid lines of the focused 8192-pulse raw
    cube), complex64 planes resident on the device. A step is what
    insar_process does on the device: the azimuth shift and flattening of every
    channel (osar_insar_coregister) and the 5 x 5 Hann multilook of the
    intensities, the six interferograms and their coherences over the region
    of interest of the 1024 m scene (osar_insar_multilook). value = SLC
    samples (channel x line x range sample) per second."""
    import torch
    from oceansar_amd import _lib
    from oceansar_amd.insar_processor import CoregPlan, coregister_channel, multilook
    nch, naz, nrg = 4, 7555, 624
    rng = np.random.default_rng(9)
    x = (rng.standard_normal((nch, 1, naz, nrg), dtype=np.float32) +
         1j * rng.standard_normal((nch, 1, naz, nrg), dtype=np.float32)).astype(np.complex64)
    dev = torch.device('cuda', 0)
    slc = torch.as_tensor(x, device=dev)
    ctx = _lib.context(0)
    v, prf = 6900., 1800.
    # the host tables of each channel (kernel spectrum, flat-earth phasor), built once
    plans = [CoregPlan(naz, -b / (v / prf), np.exp(1j * 1e-3 * b * np.arange(nrg)), dev)
             for b in (0., 5., 10., 20.)]
    crop = (naz // 2 - 133, naz // 2 + 134, 0, 431)
    images = torch.empty((nch, naz, nrg), dtype=torch.complex128, device=dev)
    ev = []

    def step(record):
        for ch in range(nch):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            coregister_channel(ctx, slc[ch], plans[ch], out=images[ch:ch + 1])
            b.record()
            if record:
                ev.append(('coregister', a, b))
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        multilook(ctx, images, 5, 5, 'hanning', crop)
        b.record()
        if record:
            ev.append(('multilook', a, b))

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    progress('insar steps done: %.3f ms per step' % (dt * 1e3))
    per = {}
    co = []
    for k, a, b in ev:
        ms = a.elapsed_time(b)
        per[k] = per.get(k, 0.) + ms / args.steps
        if k == 'coregister':
            co.append(ms)
    samples = nch * naz * nrg
    co_bytes = naz * nrg * (8 + 16)          # complex64 in, complex128 out per channel launch
    cpu = None
    if not args.no_cpu:
        # the numpy restatement of the same step (oracle/insar.py), on the host cores
        import oracle.insar as oi
        threads, n_os, model = host_cores()
        t0 = time.perf_counter()
        p = oi.coregister(x.copy(), [0., 5., 10., 20.], v, prf)
        p = p * np.stack([np.exp(1j * 1e-3 * b * np.arange(nrg)) for b in (0., 5., 10., 20.)]
                         )[:, None, None, :]
        oi.l1b(p, 5, 5, 'hanning', *crop)
        t_cpu = time.perf_counter() - t0
        cpu = {'value': samples / t_cpu, 'unit': 'samples/s', 'cores': 1, 'kind': 'port',
               'sample': 'the whole step (4 x 7555 x 624), numpy complex128 FFTs and scipy '
                         'fftconvolve as the reference; %.1f s' % t_cpu,
               'host': {'os_cpu_count': n_os, 'cpu_model': model}}
        progress('insar cpu baseline done (%.1f s)' % t_cpu)
    achieved = co_bytes / (np.mean(co) * 1e-3) / 1e9
    return {
        'metric': 'SLC samples/s (InSAR L1b processing: azimuth shift, flattening, multilook, '
                  'coherence)',
        'value': samples / dt, 'unit': 'samples/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': dt * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64 (complex64 SLC in)',
        'data': 'synthetic (seeded complex64 SLC planes)',
        'config': {'workload': 'InSAR L1b of the large workload: 4 ch x 7555 lines x 624 samples, '
                               '5x5 Hann multilook over 267 x 431', 'stage_ms_per_step': per},
        'roofline': {'bound': 'hbm', 'kernel': 'coreg_kernel', 'achieved': achieved,
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                     'traffic': None, 'avg_launch_ms': float(np.mean(co)),
                     'algorithmic_bytes_per_launch': co_bytes,
                     'note': 'per channel launch: 8 B complex64 read + 16 B complex128 written '
                             'per SLC sample; four 8192-point FP64 FFTs per column in LDS'},
        'cpu_baseline': cpu,
    }


def bench_focus(args):
    """SAR focusing of the large workload's raw cube (sar_focus's processing
    loop on the device, osar_sar_focus): 4 channels x 8192 pulses x 624 range
    samples, complex64 raw planes resident on the device, padded to 8192 x 1024
    and focused in FP64 (four in-LDS FFT passes over a complex128 workspace).
    value = raw samples (channel x pulse x range sample) focused per second."""
    import ctypes
    import torch
    from oceansar_amd import _lib
    from oceansar_amd.sar_processor import FocusParams
    nch, naz0, nrg0 = 4, 8192, 624
    n_az, n_rg = 8192, 1024
    rng = np.random.default_rng(7)
    x = (rng.standard_normal((1, nch, naz0, nrg0), dtype=np.float32) +
         1j * rng.standard_normal((1, nch, naz0, nrg0), dtype=np.float32)).astype(np.complex64)
    dev = torch.device('cuda', 0)
    raw = torch.as_tensor(x, device=dev)
    # C band, 1800 Hz PRF, 110 MHz range sampling, 1.528 m antenna (the large workload)
    params = FocusParams(1800., 110e6, 5.405e9, 6900., 616400.5, 1100., 0.6, ant_l=1.528)
    fr, rcmc, h = params.tables(n_az, n_rg)
    lo, hi = params.valid_rows(naz0)
    fr_d = torch.as_tensor(fr, dtype=torch.float64, device=dev)
    rcmc_d = torch.as_tensor(rcmc, dtype=torch.float64, device=dev)
    h_d = torch.as_tensor(h, dtype=torch.complex128, device=dev)
    work = torch.empty(n_az * n_rg, dtype=torch.complex128, device=dev)
    slc = torch.empty((nch, 1, hi - lo, nrg0), dtype=torch.complex128, device=dev)
    ctx = _lib.context(0)
    descs = []
    for ch in range(nch):
        d = _lib.FocusDesc()
        d.n_pol, d.n_az0, d.n_rg0, d.n_az, d.n_rg = 1, naz0, nrg0, n_az, n_rg
        d.az_lo, d.az_hi = lo, hi
        d.raw_pol_stride = raw.stride(0)
        d.raw = raw[0, ch].data_ptr()
        d.fr, d.rcmc, d.h = fr_d.data_ptr(), rcmc_d.data_ptr(), h_d.data_ptr()
        d.work, d.slc = work.data_ptr(), slc[ch].data_ptr()
        descs.append(d)
    ev = []

    def step(record):
        for d in descs:
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            _lib.check(ctx.lib.osar_sar_focus(ctx.h, ctypes.byref(d), _lib.stream_ptr()))
            b.record()
            if record:
                ev.append((a, b))

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    progress('focus steps done: %.3f ms per step' % (dt * 1e3))
    ms = [a.elapsed_time(b) for a, b in ev]
    samples = nch * naz0 * nrg0
    # per channel call: complex64 raw read, the complex128 workspace written by
    # the range pass, read and written by the two middle passes, read by the
    # last one, and the kept complex128 SLC written
    nbytes = (8 * naz0 * nrg0 + 16 * n_az * n_rg * 5 + 16 * (hi - lo) * nrg0)
    cpu = None
    if not args.no_cpu:
        import oracle.focus as of
        threads, n_os, model = host_cores()
        t0 = time.perf_counter()
        of.focus_channel(x[:, 0], 1800., 110e6, 5.405e9, 6900., 616400.5, 1100., 0.6,
                         None, None, ant_l=1.528)
        t_cpu = time.perf_counter() - t0
        cpu = {'value': naz0 * nrg0 / t_cpu, 'unit': 'samples/s', 'cores': 1, 'kind': 'port',
               'sample': 'one channel (8192 x 624), numpy complex128 FFTs as the reference; '
                         '%.1f s' % t_cpu, 'host': {'os_cpu_count': n_os, 'cpu_model': model}}
        progress('focus cpu baseline done (%.1f s)' % t_cpu)
    achieved = nbytes / (np.mean(ms) * 1e-3) / 1e9
    return {
        'metric': 'raw samples/s (SAR focusing: range/azimuth FFTs, RCMC, azimuth compression)',
        'value': samples / dt, 'unit': 'samples/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': dt * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64 (complex64 raw in)',
        'data': 'synthetic (seeded complex64 raw planes)',
        'config': {'workload': 'sar_focus of the large workload: 4 ch x 8192 pulses x 624 '
                               'samples, padded to 8192 x 1024, %d lines kept' % (hi - lo)},
        'roofline': {'bound': 'hbm', 'kernel': 'osar_sar_focus (4 passes)', 'achieved': achieved,
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                     'traffic': None, 'avg_launch_ms': float(np.mean(ms)),
                     'algorithmic_bytes_per_launch': nbytes,
                     'note': 'per channel call (four kernel launches): 8 B raw read, five '
                             'complex128 workspace sweeps, the kept SLC written'},
        'cpu_baseline': cpu,
    }


def bench_init(args):
    """Surface initialisation on the device (osar_surface_init): the Elfouhaily
    spectrum, Elfouhaily spreading and Philox draws of the large workload's
    2048 x 2048 surface (0.5 m). value = grid points per second."""
    import torch
    from oceansar_amd.surface_init import init_surface, init_surface_device
    kw = dict(Lx=1024., Ly=1024., dx=0.5, dy=0.5, spec_model='elfouhaily',
              spread_model='elfouhaily', wind_dir=0., wind_fetch=500e3, wind_U=12.,
              current_mag=0., current_dir=0.)
    ms = []
    for i in range(args.warmup + args.steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        s = init_surface_device(seed=11 + i, **kw)
        b.record()
        torch.cuda.synchronize()
        if i >= args.warmup:
            ms.append(a.elapsed_time(b))
    n = s.Nx * s.Ny
    dt = float(np.mean(ms)) * 1e-3
    progress('device init done: %.3f ms per surface' % (dt * 1e3))
    nbytes = n * (8 + 4)           # complex64 coefficients and float32 spectrum written
    cpu = None
    if not args.no_cpu:
        threads, n_os, model = host_cores()
        t0 = time.perf_counter()
        init_surface(rng_wave=np.random.default_rng(3), **kw)
        t_cpu = time.perf_counter() - t0
        cpu = {'value': n / t_cpu, 'unit': 'points/s', 'cores': 1, 'kind': 'port',
               'sample': 'the same 2048 x 2048 surface, host numpy init (float32 tables '
                         'as the reference); %.1f s' % t_cpu,
               'host': {'os_cpu_count': n_os, 'cpu_model': model}}
    achieved = nbytes / dt / 1e9
    return {
        'metric': 'grid points/s (OceanSurface.init on the device: spectrum, spreading, '
                  'random coefficients)',
        'value': n / dt, 'unit': 'points/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': dt * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64 (float32 tables out)',
        'data': 'synthetic (Philox4x32-10 draws)',
        'config': {'workload': '2048 x 2048 surface at 0.5 m, U10 12 m/s, Elfouhaily spectrum '
                               'and spreading (host call included: the host computes the '
                               'grid and scalars, the device the %d points)' % n},
        'roofline': {'bound': 'valu', 'kernel': 'surface_init_kernel', 'achieved': achieved,
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                     'traffic': None, 'algorithmic_bytes_per_launch': nbytes,
                     'note': 'FP64 spectrum and spreading per point; 12 B written per point'},
        'cpu_baseline': cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--workload', default='large', choices=configs_names())
    ap.add_argument('--pulses', type=int, default=None, help='shorter scene (testing only)')
    ap.add_argument('--chunk', type=int, default=512,
                    help='snapshot intervals per launch group (512: the 8192-pulse run of '
                         'the 2048^2 workload in one group, about 164 GB of fields, synthesis '
                         'workspace and amplitudes on the device; 128 / 256 measured 2.1 / '
                         '0.9 %% slower, with 41 / 82 GB)')
    ap.add_argument('--cpu-sample', type=int, default=96,
                    help='pulses of the all-core CPU sample (also the accuracy window)')
    ap.add_argument('--cpu-sample-1t', type=int, default=8,
                    help='pulses of the one-thread CPU sample (C loop only)')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-fp64', action='store_true',
                    help='skip the all-FP64 GPU run the timed raw cube is compared with')
    ap.add_argument('--no-serve', action='store_true',
                    help='skip the serving run whose p50/p99 latencies join the default line')
    ap.add_argument('--overlap', action='store_true',
                    help='two streams: synthesis/Bragg of chunk c+1 beside the accumulation '
                         'of chunk c (measured no faster on the large workload: every stage '
                         'already fills the GPU)')
    ap.add_argument('--form', default='plain',
                    choices=('plain', 'batch', 'serve', 'insar', 'focus', 'init'),
                    help='plain: one scene (the headline line); batch: 64 mixed scenes, '
                         'plain vs concurrent vs one graph launch; serve: Poisson job '
                         'arrivals, p50/p99 latency vs sequential plain execution; insar: '
                         'InSAR L1b processing of the large workload\'s 4-channel SLC; '
                         'focus: sar_focus of its raw cube; init: OceanSurface.init of its '
                         'surface on the device')
    ap.add_argument('--jobs', type=int, default=None, help='scenes (batch) or jobs (serve)')
    ap.add_argument('--batch-chunk', type=int, default=64,
                    help='batch form: snapshot intervals per persistent round (64: 4 rounds '
                         'on the 64-scene mix, 1.01x the plain path; 32: 0.90x, 128: 1.00x)')
    ap.add_argument('--streams', type=int, default=4)
    ap.add_argument('--host-init', action='store_true',
                    help='serve form: build the job surfaces on the host instead of the device')
    ap.add_argument('--copy-probe-mib', type=int, default=0,
                    help='after timing, run one device copy of this many MiB (a known-size '
                         'kernel that calibrates the FETCH_SIZE/WRITE_SIZE counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: this build has no multi-GPU path '
              '(pulse sharding across GPUs is a design note, not implemented)', file=sys.stderr)
        sys.exit(2)

    if args.form == 'batch':
        print(json.dumps(bench_batch(args)))
        return
    if args.form == 'serve':
        print(json.dumps(bench_serve(args)))
        return
    if args.form == 'insar':
        print(json.dumps(bench_insar(args)))
        return
    if args.form == 'focus':
        print(json.dumps(bench_focus(args)))
        return
    if args.form == 'init':
        print(json.dumps(bench_init(args)))
        return

    import torch
    from oceansar_amd import configs
    from oceansar_amd.engine import DeviceScene

    over = {} if args.pulses is None else {'n_pulses': args.pulses}
    sc = configs.load_scene(configs.workload(args.workload, **over), tempfile.mkdtemp())
    ds = DeviceScene(sc)
    W = sc.contributions
    raw = ds.new_raw()
    overlap = args.overlap
    # the launch group's fields, amplitudes and synthesis workspace must fit
    # the free device memory: halve it until they take at most 90 % of it
    from oceansar_amd.batch import fit_chunk
    chunk = fit_chunk(sc, args.chunk, torch.cuda.mem_get_info()[0], nbuf=2 if overlap else 1)
    for _ in range(args.warmup):
        ds.zero_c32(raw)
        ds.run(raw=raw, chunk_snaps=chunk, overlap=overlap)
    torch.cuda.synchronize()
    progress('warm-up done (%d steps, launch group %d intervals)' % (args.warmup, chunk))

    stage_events = []

    def timer(stage, a, b):
        stage_events.append((stage, a, b))

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ds.zero_c32(raw)
        ds.run(raw=raw, chunk_snaps=chunk, timer=timer, overlap=overlap)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    progress('timed steps done: %.1f ms per step' % (dt * 1e3))

    # the same steps ending with the raw cube on the host (pinned memory), the
    # scope "first launch to raw data on the host"
    host_raw = torch.empty(raw.shape, dtype=raw.dtype, pin_memory=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ds.zero_c32(raw)
        ds.run(raw=raw, chunk_snaps=chunk, overlap=overlap)
        host_raw.copy_(raw, non_blocking=True)
        torch.cuda.synchronize()
    dt_host = (time.perf_counter() - t0) / args.steps
    progress('steps with the host copy done: %.1f ms per step' % (dt_host * 1e3))

    per_stage = {}
    echo_ms = []
    for stage, a, b in stage_events:
        ms = a.elapsed_time(b)
        per_stage[stage] = per_stage.get(stage, 0.) + ms / args.steps
        if stage == 'echo':
            echo_ms.append(ms)
    npol = len(sc.pols)
    fpc = flop_per_contribution(sc.num_ch, npol)
    n_launch = len(echo_ms)
    contrib_per_launch = W / (n_launch / args.steps)
    achieved = contrib_per_launch * fpc / (np.mean(echo_ms) * 1e-3) / 1e12
    pulses_per_launch = int(round(sc.az_steps / (n_launch / args.steps)))
    traffic, traffic_src = measured_traffic(sc, repr(ds.shape), pulses_per_launch)

    out = {
        'metric': 'facet-pulse contributions/s (raw-data generation, rel. RMS <= 1e-4 vs FP64)',
        'value': W / dt,
        'unit': 'contributions/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': dt * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 (FP64 range/phase/bins)',
        'data': 'synthetic (seeded Elfouhaily surface, seeded scatterers)',
        'config': {'workload': '%s: %dx%d facets, %d pulses, %d channels, %s, K=%d'
                               % (args.workload, sc.Nx, sc.Ny, sc.az_steps, sc.num_ch, sc.pol,
                                  sc.K),
                   'facets': sc.Nx * sc.Ny, 'pulses': sc.az_steps, 'channels': sc.num_ch,
                   'range_bins': sc.rg_samp, 'contributions': W,
                   'stage_ms_per_step': {k: round(v, 3) for k, v in per_stage.items()},
                   'echo_launch_shape': repr(ds.shape),
                   'launch_group_intervals': chunk},
        'value_with_host_copy': W / dt_host,
        'ms_per_step_with_host_copy': dt_host * 1e3,
        'host_copy_note': 'the same steps, each ending with the %.0f MB raw cube copied to '
                          'pinned host memory' % (raw.numel() * 8 / 1e6),
        'roofline': {'bound': 'valu', 'kernel': 'echo_accumulate_kernel',
                     'achieved': achieved, 'peak': FP32_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': achieved / FP32_PEAK_TFLOPS, 'traffic': traffic,
                     'traffic_source': traffic_src,
                     'traffic_gbs': (None if traffic is None
                                     else traffic / (np.mean(echo_ms) * 1e-3) / 1e9),
                     'flop_per_contribution': fpc,
                     'pulses_per_launch': pulses_per_launch,
                     'avg_launch_ms': float(np.mean(echo_ms)),
                     'note': 'vector-ALU bound (FP32/FP64 VALU + transcendental + LDS atomics; '
                             'no MFMA); peak is the MI355X FP32 vector figure'},
    }
    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.neg(src, out=dst)      # an elementwise kernel: reads and writes n*4 bytes
        torch.cuda.synchronize()
        del src, dst
    if not args.no_cpu:
        out['cpu_baseline'], out['accuracy'] = cpu_baseline(sc, raw, args.cpu_sample,
                                                            args.cpu_sample_1t)
    if not args.no_fp64:
        # the whole timed raw cube against the all-FP64 GPU path on the same inputs
        mixed = host_raw.numpy().astype(np.complex128)
        torch.cuda.empty_cache()
        t0 = time.perf_counter()
        ref64, _ = ds.run(chunk_snaps=64, fp64=True)
        torch.cuda.synchronize()
        t_fp64 = time.perf_counter() - t0
        progress('all-FP64 GPU run done (%.1f s)' % t_fp64)
        ref64 = ref64.cpu().numpy()
        e = float(np.linalg.norm(mixed - ref64) / np.linalg.norm(ref64))
        mag = np.abs(ref64)
        keep = mag > mag.max() * 10 ** (-30. / 20.)
        ph = float(np.sqrt(np.mean(np.angle(mixed[keep] * np.conj(ref64[keep])) ** 2)))
        acc = out.setdefault('accuracy', {})
        acc['whole_run_vs_fp64_gpu'] = {
            'rel_rms': e, 'phase_rms_rad': ph,
            'against': 'the all-FP64 GPU path (FP64 synthesis, Bragg and accumulation; '
                       'pinned to the FP64 oracle at a few 1e-15) over all %d pulses and %d '
                       'channels' % (sc.az_steps, sc.num_ch),
            'fp64_run_s': t_fp64}
        del ref64, mixed
    del ds, raw, host_raw
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    if not args.no_serve:
        sv = bench_serve(args)
        out['serving'] = {k: sv[k] for k in ('served', 'plain_sequential', 'p50_speedup_vs_plain',
                                             'p99_speedup_vs_plain', 'throughput_vs_plain')}
        out['serving']['workload'] = sv['config']['workload']
        out['p50_latency_s'] = sv['served']['p50_latency_s']
        out['p99_latency_s'] = sv['served']['p99_latency_s']
    print(json.dumps(out))


if __name__ == '__main__':
    main()
