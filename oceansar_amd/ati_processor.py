"""ATI (along-track interferometry) analysis of the InSAR L1b file: ati_process().

Mirrors oceansar/radarsim/ati_processor.py:42-470 (same signature, inputs and
output files; no plots): the L1b products (:78-96), the surface velocities at
t = 0 of the ocean state file, synthesised on the GPU (:101-104), their
statistics and multilooked versions (:121-145, the smoothing through
osar_insar_multilook), the ATI phase of the first channel pair per
polarisation with the histogram-centred unwrapping uwphase (:31-39, :243-253),
the radial and horizontal velocity estimates (:366-376), the estimated mean
NRCS (:411-417), the text report with the reference's lines (:336-423) and
ati_stats.npz (:443-462).

Quirks kept: the L1b inc_angle holds radians (insar_process writes the value
after its own deg2rad) and is converted with deg2rad once more here, as the
reference reads it (:92); the region trimmed by az_guard uses Python slices
[az_guard:-az_guard], which are empty when az_guard is 0.
"""
import os

import numpy as np

from . import _lib
from . import constants as const
from . import raw_io
from .cfg import ConfigFile
from .geometry import inc_to_sr
from .insar_processor import multilook
from .ocean import OceanSurface


def uwphase(phasor):
    """Phases re-centred on the fullest histogram bin (insar_processor.py:31-39):
    at most 72 bins over [-pi, pi), the centre taken as the left edge of that
    bin (the reference averages the edge with itself)."""
    pha = np.angle(phasor)
    nbins = np.min([72, np.ceil(phasor.size / 50).astype(int)])
    hist, edges = np.histogram(pha, bins=nbins, range=(-np.pi, np.pi))
    centre = edges[hist.argmax()]
    return np.angle(phasor * np.exp(-1j * centre)) + centre


def smooth2(ctx, field, rg_len, az_len):
    """utils.smooth(utils.smooth(field, rg_len, axis=1), az_len, axis=0) with
    flat windows (misc.py:108-145) on the device, FP64 sums."""
    import torch
    dev = torch.device('cuda', ctx.device)
    img = torch.as_tensor(np.ascontiguousarray(field, dtype=np.float64), device=dev)[None]
    out, _, _ = multilook(ctx, img, rg_len, az_len, 'flat')
    return out[0].cpu().numpy()


def ati_process(cfg_file, insar_output_file, ocean_file, output_file, device=0):
    """ati_processor.py:42-470. Returns the dictionary of the values written
    to the report and to ati_stats.npz."""
    cfg = ConfigFile(cfg_file)
    pol = cfg.sar.pol
    polt = ['hh', 'vv'] if pol == 'DP' else (['hh'] if pol == 'hh' else ['vv'])
    with raw_io.L1bFile(insar_output_file, 'r') as lf:
        i_all = lf.get('ml_intensity')
        coh_lut = lf.get('coh_lut')
        if lf._f.dimensions['covel_dim']:
            cohs = lf.get('ml_coherence') * np.exp(1j * lf.get('ml_phase'))
        else:
            cohs = np.zeros((0,) + i_all.shape[2:], dtype=complex)
        b_ati = np.atleast_1d(lf.get('b_ati'))
        f0 = lf.get('f0')
        az_sampling = lf.get('az_sampling')
        num_ch = int(lf.get('num_ch'))
        rg_sampling = lf.get('rg_sampling')
        v_ground = lf.get('v_ground')
        alt = lf.get('orbit_alt')
        inc_angle = np.deg2rad(lf.get('inc_angle'))
        rg_ml = int(lf.get('rg_ml'))
        az_ml = int(lf.get('az_ml'))
    k0 = 2. * np.pi * f0 / const.c

    surface = OceanSurface(device)
    surface.load(ocean_file, compute=['D', 'V'])
    surface.t = 0.
    ctx = _lib.context(device)

    grg_grid_spacing = const.c / 2. / rg_sampling / np.sin(inc_angle)
    az_grid_spacing = v_ground / az_sampling
    res_fact = np.ceil(np.sqrt((grg_grid_spacing / surface.dx) * (az_grid_spacing / surface.dy)))
    # float32 fields times float64 scalars stay float32 (numpy < 1.24 casting)
    v_r = (surface.Vx * np.float32(np.sin(inc_angle)) -
           surface.Vz * np.float32(np.cos(inc_angle)))
    v_r_ml = smooth2(ctx, v_r, res_fact * rg_ml, res_fact * az_ml)
    v_h = surface.Vx
    v_h_ml = smooth2(ctx, v_h, res_fact * rg_ml, res_fact * az_ml)
    st = {'v_radial_surf_mean': np.mean(v_r), 'v_radial_surf_std': np.std(v_r),
          'v_radial_surf_ml_std': np.std(v_r_ml), 'v_horizo_surf_mean': np.mean(v_h),
          'v_horizo_surf_std': np.std(v_h), 'v_horizo_surf_ml_std': np.std(v_h_ml),
          'grg_grid_spacing': grg_grid_spacing, 'az_grid_spacing': az_grid_spacing}
    sr0 = inc_to_sr(inc_angle, alt)
    std_az_shift = st['v_radial_surf_std'] / v_ground * sr0
    az_guard = int(std_az_shift / (v_ground / az_sampling))
    st['az_guard'] = az_guard
    npol = i_all.shape[1]
    tau_ati = b_ati / v_ground
    g = slice(az_guard, -az_guard)
    out = open(output_file, 'w')
    if num_ch > 1:
        phases = np.array([uwphase(cohs[coh_lut[(0, p, 1, p)]]) for p in range(npol)])
        v_rad = (-phases / tau_ati[1] / (k0 * 2.))[:, g, 5:-5]
        v_hor = (-phases / tau_ati[1] / (k0 * 2.) / np.sin(inc_angle))[:, g, 5:-5]
        sm = [smooth2(ctx, v_rad[p], rg_ml, az_ml) if v_rad[p].size else v_rad[p]
              for p in range(npol)]
        st['v_radial_est_mean'] = [np.mean(v_rad[p]) for p in range(npol)]
        st['v_radial_est_std'] = [np.std(v_rad[p]) for p in range(npol)]
        st['v_radial_est_ml_mean'] = [np.mean(s) for s in sm]
        st['v_radial_est_ml_std'] = [np.std(s) for s in sm]
        st['v_horizo_est_mean'] = [np.mean(v_hor[p]) for p in range(npol)]
        st['v_horizo_est_std'] = [np.std(v_hor[p]) for p in range(npol)]
        sep = '--------------------------------------------\n'
        w = out.write
        w(sep + 'SURFACE RADIAL VELOCITY - NO SMOOTHING\n')
        w('MEAN(SURF. V) = %.4f\n' % st['v_radial_surf_mean'])
        w('STD(SURF. V) = %.4f\n' % st['v_radial_surf_std'])
        w(sep + '\n')
        w(sep + 'SURFACE RADIAL VELOCITY - SMOOTHING (WIN. SIZE=%dx%d)\n' % (az_ml, rg_ml))
        w('MEAN(SURF. V) = %.4f\n' % st['v_radial_surf_mean'])
        w('STD(SURF. V) = %.4f\n' % st['v_radial_surf_ml_std'])
        w(sep + '\n')
        w(sep + 'SURFACE HORIZONTAL VELOCITY - NO SMOOTHING\n')
        w('MEAN(SURF. V) = %.4f\n' % st['v_horizo_surf_mean'])
        w('STD(SURF. V) = %.4f\n' % st['v_horizo_surf_std'])
        w(sep + '\n')
        w(sep + 'ESTIMATED RADIAL VELOCITY - NO SMOOTHING\n')
        for p in range(npol):
            w('%s Polarization\n' % polt[p])
            w('MEAN(EST. V) = %.4f\n' % st['v_radial_est_mean'][p])
            w('STD(EST. V) = %.4f\n' % st['v_radial_est_std'][p])
        w(sep + '\n')
        w(sep + 'ESTIMATED RADIAL VELOCITY - SMOOTHING (WIN. SIZE=%dx%d)\n' % (az_ml, rg_ml))
        for p in range(npol):
            w('%s Polarization\n' % polt[p])
            w('MEAN(EST. V) = %.4f\n' % st['v_radial_est_ml_mean'][p])
            w('STD(EST. V) = %.4f\n' % st['v_radial_est_ml_std'][p])
        w(sep + '\n')
        w(sep + 'ESTIMATED HORIZONTAL VELOCITY - NO SMOOTHING\n')
        for p in range(npol):
            w('%s Polarization\n' % polt[p])
            w('MEAN(EST. V) = %.4f\n' % st['v_horizo_est_mean'][p])
            w('STD(EST. V) = %.4f\n' % st['v_horizo_est_std'][p])
        w(sep + '\n')
    nrcs = 10 * np.log10(np.mean(np.mean(i_all[:, :, g, 5:-5], axis=-1), axis=-1))
    st['nrcs'] = nrcs
    sep = '--------------------------------------------\n'
    out.write(sep)
    for p in range(npol):
        out.write('%s Polarization\n' % polt[p])
        out.write('Estimated mean NRCS = %5.2f\n' % nrcs[0, p])
    out.write(sep + '\n')
    out.write(sep)
    out.write('GROUND RANGE GRID SPACING = %.4f\n' % grg_grid_spacing)
    out.write('AZIMUTH GRID SPACING = %.4f\n' % az_grid_spacing)
    out.write(sep + '\n')
    out.close()
    if num_ch > 1:
        c = np.array(cohs)[:, g, 5:-5]
        st['v_r_dop'] = np.mean(np.mean(v_rad, axis=-1), axis=-1)
        st['coh_mean'] = np.mean(np.mean(c, axis=-1), axis=-1)
        st['abscoh_mean'] = np.mean(np.mean(np.abs(c), axis=-1), axis=-1)
        np.savez(os.path.join(os.path.dirname(output_file), 'ati_stats.npz'),
                 nrcs=nrcs, v_r_dop=st['v_r_dop'], v_r_surf=st['v_radial_surf_mean'],
                 v_r_surf_std=st['v_radial_surf_std'], coh_mean=st['coh_mean'],
                 abscoh_mean=st['abscoh_mean'], coh_lut=coh_lut, pols=polt)
    return st
