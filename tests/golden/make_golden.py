"""Write golden input/output vectors of the reference's pure-numpy leaf functions.

Run in a container where the reference checkout exists (read-only):
    python tests/golden/make_golden.py /root/reference

Only the functions that no closed form pins bit-for-bit are exercised here:
spec.elfouhaily (its branch constants), utils.geometry.gr_to_geo (spline
details) and inc_to_sr / inc_to_gr / orbit_to_vel, utils.misc.optimize_fftsize,
closure.grid_coherence, radarsim.antenna.sinc_bp,
radarsim.range_profile.calc_sinc_vec and swell_spec.ardhuin_swell_spec.
Each source file is read as text and executed in a fresh namespace (never
through the import system, so no bytecode is read or written), with the
reference's own sibling modules supplied the same way. The output is data
only: golden.npz (inputs and outputs), which the tests load with
allow_pickle=False.
"""
import os
import sys
import types

import numpy as np


def _exec_source(path, name, extra=None):
    mod = types.ModuleType(name)
    mod.__file__ = path
    if extra:
        mod.__dict__.update(extra)
    with open(path, 'r') as fh:
        src = fh.read()
    exec(compile(src, path, 'exec'), mod.__dict__)
    return mod


def _function_source(path, func_name):
    """Source text of one top-level function (for modules whose other imports are absent)."""
    with open(path) as fh:
        lines = fh.read().split('\n')
    out, grab = [], False
    for ln in lines:
        if ln.startswith('def ' + func_name + '('):
            grab = True
        elif grab and ln and not ln[0].isspace() and not ln.startswith('#'):
            break
        if grab:
            out.append(ln)
    return '\n'.join(out)


def main(ref):
    sys.dont_write_bytecode = True
    pkg = os.path.join(ref, 'oceansar')
    const = _exec_source(os.path.join(pkg, 'constants', 'constants.py'), 'oceansar.constants')
    # a minimal 'oceansar' namespace so that 'from oceansar import constants' resolves
    oceansar = types.ModuleType('oceansar')
    oceansar.constants = const
    sys.modules['oceansar'] = oceansar
    sys.modules['oceansar.constants'] = const
    misc = _exec_source(os.path.join(pkg, 'utils', 'misc.py'), 'oceansar.utils.misc')
    utils = types.ModuleType('oceansar.utils')
    utils.optimize_fftsize = misc.optimize_fftsize
    oceansar.utils = utils
    sys.modules['oceansar.utils'] = utils
    geom = _exec_source(os.path.join(pkg, 'utils', 'geometry.py'), 'oceansar.utils.geometry')
    spec = _exec_source(os.path.join(pkg, 'spec', 'elfouhaily.py'), 'oceansar.spec.elfouhaily')
    coh = _exec_source(os.path.join(pkg, 'closure', 'coherence_times.py'),
                       'oceansar.closure.coherence_times')
    ant = _exec_source(os.path.join(pkg, 'radarsim', 'antenna.py'), 'oceansar.radarsim.antenna')
    swell = _exec_source(os.path.join(pkg, 'swell_spec', 'dir_swell_spec.py'),
                         'oceansar.swell_spec.dir_swell_spec')
    # range_profile.py imports numba at module level; calc_sinc_vec itself is numpy only
    rp = types.ModuleType('rp')
    exec(compile('import numpy as np\n' +
                 _function_source(os.path.join(pkg, 'radarsim', 'range_profile.py'),
                                  'calc_sinc_vec'), 'range_profile.py', 'exec'), rp.__dict__)

    g = {}
    # spectrum at the wind speeds of the workloads, float64 and float32 inputs
    k64 = np.concatenate([np.geomspace(0.005, 230., 300), np.arange(1, 2 * 113.28, 0.1)[::37]])
    g['spec_k'] = k64
    for U in (3., 7., 10., 12., 15.):
        g['spec_U%g' % U] = spec.elfouhaily(k64, U, 500e3)
        g['spec32_U%g' % U] = spec.elfouhaily(k64.astype(np.float32), np.float32(U), 500e3)
    # geometry
    alt = 514e3
    inc = np.deg2rad(35.)
    g['geo_sr0'] = np.array([geom.inc_to_sr(inc, alt)])
    g['geo_gr0'] = np.array([geom.inc_to_gr(inc, alt)])
    x = np.linspace(-512., 512., 333, dtype=np.float32)
    gr = x + g['geo_gr0'][0]
    sr, th_i, th_l = geom.gr_to_geo(gr, alt)
    g['geo_gr'] = gr
    g['geo_sr'] = sr
    g['geo_inc'] = th_i
    g['geo_look'] = th_l
    g['geo_v_ground'] = np.array([geom.orbit_to_vel(alt, ground=True), geom.orbit_to_vel(693e3)])
    # sizing
    vals = np.array([115, 267, 624, 1000, 37736, 53348, 513, 2049, 300])
    g['fft_in'] = vals
    g['fft_p2'] = np.array([misc.optimize_fftsize(int(v)) for v in vals])
    g['fft_p3'] = np.array([misc.optimize_fftsize(int(v), 3) for v in vals])
    g['fft_p5'] = np.array([misc.optimize_fftsize(int(v), 5) for v in vals])
    # closure
    Us = np.array([3., 7., 10., 12., 15.])
    res = np.array([0.5, 1.0])
    g['tau_U'] = Us
    g['tau_res'] = res
    g['tau'] = np.array([[coh.grid_coherence(U, r, 5.405e9) for r in res] for U in Us])
    # antenna
    s = np.linspace(-0.03, 0.03, 201)
    g['bp_sin'] = s
    g['bp_field'] = ant.sinc_bp(s, 1.528, 5.405e9, field=True)
    g['bp_power'] = ant.sinc_bp(s, 12.6, 5.405e9)
    # range response table
    g['sinc_vec_fs1p1'] = rp.calc_sinc_vec(8, 20, Fs=1.1)
    g['sinc_vec_fs1'] = rp.calc_sinc_vec(6, 10, Fs=1.0)
    # swell
    kk = np.geomspace(0.005, 2., 64)
    th = np.linspace(-np.pi, np.pi, 64)
    K_, T_ = np.meshgrid(kk, th)
    g['swell_k'] = K_.ravel()
    g['swell_th'] = T_.ravel()
    g['swell'] = swell.ardhuin_swell_spec(K_.ravel(), T_.ravel(), 45., 0.0884, 0.007, 8, 2.)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden.npz')
    np.savez(out, **g)
    print('wrote', out, len(g), 'arrays')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
