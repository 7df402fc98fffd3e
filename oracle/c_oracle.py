"""ctypes wrapper of the C restatement (oracle/c/oracle_echo.c). Test infrastructure only.

simulate_c() has the same contract as oracle.sar.simulate(): the raw cube of
pulses [p0, p1) in complex128 plus NRCS_avg, with snapshots from the numpy
oracle's FP64 synthesis and Bragg RCS from oracle.bragg.
"""
import ctypes
import os
import subprocess
from ctypes import c_double, c_int32, c_int64, c_void_p

import numpy as np

from . import sar as osar

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'c', 'oracle_echo.c')
LIB = os.path.join(HERE, 'liboracle_echo.so')
MAXCH = 8


def build():
    """gcc -O3 -fopenmp -ffp-contract=off (the numpy oracle's rounding order)."""
    cmd = ['gcc', '-O3', '-march=x86-64-v3', '-fopenmp', '-ffp-contract=off', '-fno-fast-math',
           '-fPIC', '-shared', SRC, '-o', LIB + '.tmp', '-lm']
    subprocess.run(cmd, check=True)
    os.replace(LIB + '.tmp', LIB)
    return LIB


class Desc(ctypes.Structure):
    _fields_ = [(n, c_int32) for n in ('nx', 'ny', 'K', 'snap0', 'n_snap', 'p0', 'p1', 'nbins',
                                       'n_ch', 'n_pol', 'rcs0', 'n_rcs', 'bragg', 'n_sinc', 'ovs',
                                       'pad')] + \
        [(n, c_double) for n in ('t_step', 't_span', 'v_ground', 'sr0', 'gr0', 'k0', 'min_sr',
                                 'binsize', 'ltx_l', 'lrx_l', 'inc_angle')] + \
        [('b_ati', c_double * MAXCH), ('b_xti', c_double * MAXCH)] + \
        [(n, c_void_p) for n in ('sr', 'sin_inc', 'cos_inc', 'inc', 'cap_rate', 'y', 'dx', 'dy',
                                 'dz', 'amp', 'rcs_slot', 'sm', 'sp')] + \
        [('nm', c_int64), ('np_', c_int64)] + \
        [(n, c_void_p) for n in ('ind0_m', 'ind0_p', 'sinc', 'out')]


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        lib = ctypes.CDLL(LIB)
        lib.oracle_echo.argtypes = [ctypes.POINTER(Desc)]
        lib.oracle_desc_size.restype = c_int64
        lib.oracle_set_threads.argtypes = [ctypes.c_int]
        lib.oracle_set_threads.restype = ctypes.c_int
        if lib.oracle_desc_size() != ctypes.sizeof(Desc):
            raise ImportError('oracle_desc layout mismatch')
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(c_void_p)


def set_threads(n):
    """OpenMP threads of the C pulse loop; returns the count in force."""
    return load().oracle_set_threads(int(n))


def simulate_c(sc, p0=0, p1=None, snaps=None, threads=None, timings=None):
    """C-oracle raw cube for pulses [p0, p1) (un-calibrated, noise-free) and NRCS_avg.

    timings: optional dict that receives the seconds spent in the snapshot
    synthesis ('synth_s', numpy FP64, only for snapshots not yet cached), the
    Bragg RCS refreshes ('bragg_s', numpy FP64) and the C pulse loop ('loop_s')."""
    import time
    lib = load()
    if threads is not None:
        lib.oracle_set_threads(int(threads))
    p1 = sc.az_steps if p1 is None else p1
    snaps = snaps or osar.SnapshotCache(sc)
    K = sc.K
    m0 = p0 // K
    last = p1 - 1
    m1 = last // K + (1 if last % K else 0)
    t0 = time.perf_counter()
    S = [snaps.get(m) for m in range(m0, m1 + 1)]
    t_synth = time.perf_counter() - t0
    keep = []

    def arr(x, dt):
        a = np.ascontiguousarray(x, dtype=dt)
        keep.append(a)
        return a

    dx = arr(np.stack([s['Dx'] for s in S]), np.float32)
    dy = arr(np.stack([s['Dy'] for s in S]), np.float32)
    dz = arr(np.stack([s['Dz'] for s in S]), np.float32)
    npol = len(sc.pols)
    nrcs = np.zeros((npol, p1 - p0))
    d = Desc()
    d.nx, d.ny, d.K, d.snap0, d.n_snap, d.p0, d.p1 = sc.Nx, sc.Ny, K, m0, len(S), p0, p1
    d.nbins, d.n_ch, d.n_pol = sc.rg_samp, sc.num_ch, npol
    d.n_sinc, d.ovs = sc.n_sinc, sc.sinc_ovs
    d.t_step, d.t_span, d.v_ground, d.sr0, d.gr0 = sc.t_step, sc.t_span, sc.v_ground, sc.sr0, sc.gr0
    d.k0, d.min_sr, d.binsize = sc.k0, sc.min_sr, sc.sr_res / sc.over_fs
    l0 = 299792458.0 / sc.f0
    d.ltx_l, d.lrx_l, d.inc_angle = sc.ant_l_tx / l0, sc.ant_l_rx / l0, sc.inc_angle
    for i in range(sc.num_ch):
        d.b_ati[i], d.b_xti[i] = sc.b_ati[i], sc.b_xti[i]
    d.sr, d.sin_inc = _p(arr(sc.sr, np.float64)), _p(arr(sc.sin_inc, np.float64))
    d.cos_inc, d.inc = _p(arr(sc.cos_inc, np.float64)), _p(arr(sc.inc, np.float64))
    d.y = _p(arr(sc.y64, np.float64))
    d.dx, d.dy, d.dz = _p(dx), _p(dy), _p(dz)
    d.sinc = _p(arr(sc.sinc_vec, np.float32))
    out = arr(np.zeros((npol, sc.num_ch, p1 - p0, sc.rg_samp, 2)), np.float64)
    d.out = _p(out)
    t0 = time.perf_counter()
    if sc.bragg_enable:
        d.bragg = 1
        r0, r1 = int(sc.rcs_slot[p0]), int(sc.rcs_slot[last])
        amps = []
        for r in range(r0, r1 + 1):
            rcs = osar.bragg_rcs(sc, snaps, int(sc.rcs_refresh[r]))
            amps.append(np.stack([np.sqrt(rcs[pol]) for pol in sc.pols]))
            for ip, pol in enumerate(sc.pols):
                sel = sc.rcs_slot[p0:p1] == r
                nrcs[ip, sel] = np.sum(rcs[pol]) / sc.surface_area
        d.amp = _p(arr(np.stack(amps), np.float64))
        d.rcs0, d.n_rcs = r0, len(amps)
        d.rcs_slot = _p(arr(sc.rcs_slot, np.int32))
        d.cap_rate = _p(arr(sc.cap_rate, np.float64))
        d.sm = _p(arr(sc.scat_m.s.view(np.float64), np.float64))
        d.sp = _p(arr(sc.scat_p.s.view(np.float64), np.float64))
        d.nm, d.np_ = sc.scat_m.N, sc.scat_p.N
        d.ind0_m = _p(arr(sc.scat_m.ind0, np.int64))
        d.ind0_p = _p(arr(sc.scat_p.ind0, np.int64))
    t_bragg = time.perf_counter() - t0
    t0 = time.perf_counter()
    lib.oracle_echo(ctypes.byref(d))
    t_loop = time.perf_counter() - t0
    if timings is not None:
        timings.update(synth_s=t_synth, bragg_s=t_bragg, loop_s=t_loop)
    raw = out[..., 0] + 1j * out[..., 1]
    return raw, nrcs
