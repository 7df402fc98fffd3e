"""Golden vectors of the reference's Bragg facet model, Elfouhaily spreading and
per-pulse facet range, for the functions no closed form pins in full.

Run in a container where the reference checkout exists (read-only):
    python tests/golden/make_golden_bragg.py /root/reference

What is executed, always from the reference's source text in a fresh namespace
(never through the import system, so no bytecode is read or written):
  - the body of RCSRomeiser97._rcs_numba (oceansar/nrcs/romeiser97.py:251-322),
    its two decorators (staticmethod, numba njit) dropped, so numpy evaluates the
    same statements numba compiles;
  - spread.elfouhaily (oceansar/spread/elfouhaily.py:23-56), njit dropped, with
    the module's constants rho, S, X_0 and scipy's g;
  - the two statements of sar_raw that form sr_surface and wave_dinc
    (oceansar/radarsim/sar_raw_nompi.py:366-369).
Inputs are tilted facets (slopes with sigma 0.1 plus a tail to 0.6), the
incidence and azimuth ranges of the workloads and beyond, three wind
directions. The output, golden_bragg.npz, is data only (loaded with
allow_pickle=False).
"""
import os
import re
import sys
import textwrap
import types

import numpy as np


def _exec_source(path, name):
    mod = types.ModuleType(name)
    mod.__file__ = path
    with open(path) as fh:
        exec(compile(fh.read(), path, 'exec'), mod.__dict__)
    return mod


def _def_block(lines, start):
    """Lines of the def starting at index `start` up to the next line with smaller
    or equal indentation (decorators excluded)."""
    ind = len(lines[start]) - len(lines[start].lstrip())
    out = [lines[start]]
    for ln in lines[start + 1:]:
        if ln.strip() and (len(ln) - len(ln.lstrip())) <= ind:
            break
        out.append(ln)
    return textwrap.dedent('\n'.join(out))


def _method_source(path, name):
    with open(path) as fh:
        lines = fh.read().split('\n')
    for i, ln in enumerate(lines):
        if ln.lstrip().startswith('def ' + name + '('):
            return _def_block(lines, i)
    raise KeyError(name)


def _statements(path, names):
    """Source text of the first assignment to each of `names` (in order), with
    the continuation lines of a parenthesised right-hand side."""
    with open(path) as fh:
        lines = fh.read().split('\n')
    out, i = [], 0
    for name in names:
        while not re.match(r'\s*' + name + r'\s*=', lines[i]):
            i += 1
        stmt, depth = [], 0
        while True:
            ln = lines[i].strip()
            stmt.append(ln)
            depth += ln.count('(') - ln.count(')')
            i += 1
            if depth == 0:
                break
        out.append(' '.join(stmt))
    return '\n'.join(out)


def main(ref):
    sys.dont_write_bytecode = True
    pkg = os.path.join(ref, 'oceansar')
    const = _exec_source(os.path.join(pkg, 'constants', 'constants.py'), 'oceansar.constants')

    # _rcs_numba body
    src = _method_source(os.path.join(pkg, 'nrcs', 'romeiser97.py'), '_rcs_numba')
    ns = {'np': np, 'const': const}
    exec(compile(src, 'romeiser97.py:_rcs_numba', 'exec'), ns)
    rcs_fn = ns['_rcs_numba']

    # spread.elfouhaily with its module constants
    path = os.path.join(pkg, 'spread', 'elfouhaily.py')
    with open(path) as fh:
        text = fh.read()
    consts = '\n'.join(ln for ln in text.split('\n')
                       if re.match(r'^(rho|S|X_0)\s*=', ln))
    from scipy.constants import g as g_const
    ns2 = {'np': np, 'g': g_const}
    exec(compile(consts + '\n' + _method_source(path, 'elfouhaily'), path, 'exec'), ns2)
    spread_fn = ns2['elfouhaily']

    rng = np.random.default_rng(20261204)
    k0 = 2 * np.pi * 5.405e9 / 299792458.0
    out = {'k0': np.array([k0])}
    # facets: [set][n]
    n = 1024
    sets = []
    # workload-like: 35 deg +- a few, small azimuth angles, sigma 0.1 slopes
    sets.append((np.deg2rad(rng.uniform(33., 37., n)), rng.uniform(-0.02, 0.02, n),
                 rng.normal(0., 0.1, n), rng.normal(0., 0.1, n), 0.0))
    # wider: 20-50 deg, azimuth +-0.5 rad, heavier slopes, wind at 45 deg
    sets.append((np.deg2rad(rng.uniform(20., 50., n)), rng.uniform(-0.5, 0.5, n),
                 rng.normal(0., 0.25, n), rng.normal(0., 0.25, n), np.deg2rad(45.)))
    # steep slopes (bad facets, large tilts), wind from 200 deg
    sets.append((np.deg2rad(rng.uniform(25., 45., n)), rng.uniform(-0.1, 0.1, n),
                 rng.uniform(-0.6, 1.0, n), rng.uniform(-0.6, 0.6, n), np.deg2rad(200.)))
    for i, (inc, az, sx, sy, wd) in enumerate(sets):
        k_b, th, T_hh, T_vv, bad = rcs_fn(inc, az, sx, sy, k0, 'DP', wd, 0.125)
        mask = np.zeros(n, dtype=bool)
        mask[bad] = True
        out['rcs%d_inc' % i] = inc
        out['rcs%d_az' % i] = az
        out['rcs%d_diffx' % i] = sx
        out['rcs%d_diffy' % i] = sy
        out['rcs%d_wind_dir' % i] = np.array([wd])
        out['rcs%d_k_b' % i] = k_b
        out['rcs%d_theta' % i] = th
        out['rcs%d_T_hh' % i] = T_hh
        out['rcs%d_T_vv' % i] = T_vv
        out['rcs%d_bad' % i] = mask
        # single-polarisation calls return the requested T twice
        k_v, _, Tv1, Tv2, _ = rcs_fn(inc, az, sx, sy, k0, 'vv', wd, 0.125)
        k_h, _, Th1, Th2, _ = rcs_fn(inc, az, sx, sy, k0, 'hh', wd, 0.125)
        assert np.array_equal(Tv1, T_vv) and np.array_equal(Th1, T_hh)
        assert np.array_equal(Tv2, T_vv) and np.array_equal(Th2, T_hh)

    # spreading: Bragg wave numbers of C band at 15-60 deg, angles over the circle
    kb = 2 * k0 * np.sin(np.deg2rad(np.linspace(15., 60., 41)))
    th = np.linspace(-np.pi, np.pi, 73)
    K_, T_ = np.meshgrid(kb, th)
    out['spread_k'] = K_.ravel()
    out['spread_th'] = T_.ravel()
    for U in (3., 7., 10., 12., 15.):
        for fetch in (500e3, 50e3):
            out['spread_U%g_F%g' % (U, fetch)] = spread_fn(K_.ravel(), T_.ravel(), U, fetch)

    # sr_surface and wave_dinc (sar_raw_nompi.py:366-369)
    stmt = _statements(os.path.join(pkg, 'radarsim', 'sar_raw_nompi.py'),
                       ('sr_surface', 'wave_dinc'))
    ny, nx = 16, 24
    surface = types.SimpleNamespace(Dx=rng.normal(0, 1.5, (ny, nx)).astype(np.float32),
                                    Dy=rng.normal(0, 1.5, (ny, nx)).astype(np.float32),
                                    Dz=rng.normal(0, 1.0, (ny, nx)).astype(np.float32))
    sr0 = 616400.5
    sr = np.linspace(0., 147., nx).reshape(1, nx)
    inc = np.deg2rad(np.linspace(34.9, 35.1, nx)).reshape(1, nx)
    az = np.linspace(-900., 700., ny).reshape(ny, 1)
    ns3 = {'np': np, 'surface': surface, 'sr': sr, 'cos_inc': np.cos(inc),
           'sin_inc': np.sin(inc), 'az': az, 'sin_az': az / sr0, 'sr0': sr0}
    exec(compile(stmt, 'sar_raw_nompi.py:366-369', 'exec'), ns3)
    for k in ('Dx', 'Dy', 'Dz'):
        out['geo_' + k] = getattr(surface, k)
    out['geo_sr'] = sr
    out['geo_inc'] = inc
    out['geo_az'] = az
    out['geo_sr0'] = np.array([sr0])
    out['geo_sr_surface'] = ns3['sr_surface']
    out['geo_wave_dinc'] = ns3['wave_dinc']

    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden_bragg.npz')
    np.savez(dst, **out)
    print('wrote', dst, len(out), 'arrays')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
