"""The ``.cfg``-driven raw-data generator: sar_raw() and the sarsim() entry.

sar_raw mirrors oceansar/radarsim/sar_raw_nompi.py:32-641 (same signature,
same side effects: ocean state file, NetCDF raw file). The set-up runs on the
host (scene.SarScene), the pulse loop on the GPU (engine.DeviceScene.run),
then the raw cube is trimmed to rg_samp_orig, calibrated and given complex
noise (sar_raw_nompi.py:552-574; the noise is drawn from the seeded noise
stream) and written with the reference's RawFile layout.

sarsim mirrors oceansar/radarsim/oceansar_sarsim.py:29-83: the raw stage,
then, as the cfg's [sim] flags ask, SAR focusing (sar_processor.sar_focus),
InSAR L1b processing (insar_processor.insar_process) and ATI analysis
(ati_processor.ati_process). The L2 wave-spectrum stage is not built; its
flag is reported and skipped.
"""
import os
import time

import numpy as np

from . import raw_io
from .cfg import ConfigFile
from .utils import db2lin


def _state_surface(path):
    from .ocean import OceanSurface
    s = OceanSurface()
    s.load(path, [])
    return s.state


def sar_raw(cfg_file, output_file, ocean_file, reuse_ocean_file, errors_file,
            reuse_errors_file, plot_save=True, device=0, chunk_snaps=32, verbose=True):
    """Generate raw data for cfg_file and write output_file (NetCDF RawFile).

    Returns a dict of stage timings (seconds); the reference returns None.
    errors_file / reuse_errors_file are accepted for signature compatibility
    (the reference's error model is commented out, sar_raw_nompi.py:576-581).
    plot_save: PNG plots are not produced (no plotting dependency here).
    """
    import torch
    from .engine import DeviceScene
    from .scene import SarScene

    timings = {}
    t0 = time.perf_counter()
    cfg = ConfigFile(cfg_file)
    surface = None
    if reuse_ocean_file and ocean_file and os.path.exists(ocean_file):
        try:
            surface = _state_surface(ocean_file)
        except RuntimeError:
            surface = None
    sc = SarScene(cfg, surface=surface)
    if surface is None and ocean_file:
        from .ocean import OceanSurface
        os_ = OceanSurface()
        os_._adopt(sc.surface, [])
        os_.save(ocean_file)
    timings['setup'] = time.perf_counter() - t0

    t1 = time.perf_counter()
    ds = DeviceScene(sc, device=device)
    torch.cuda.synchronize()
    timings['upload'] = time.perf_counter() - t1
    t2 = time.perf_counter()
    raw_dev, nrcs = ds.run(chunk_snaps=chunk_snaps)
    torch.cuda.synchronize()
    timings['pulse_loop'] = time.perf_counter() - t2
    t3 = time.perf_counter()
    raw = raw_dev[..., :sc.rg_samp_orig].cpu().numpy().astype(np.complex128)
    timings['download'] = time.perf_counter() - t3

    raw = calibrate_and_add_noise(sc, raw)
    t4 = time.perf_counter()
    write_raw_file(sc, output_file, raw, nrcs)
    timings['write'] = time.perf_counter() - t4
    if verbose:
        w = sc.contributions
        print('sar_raw: %dx%d facets, %d pulses, %d ch, %s: pulse loop %.3f s '
              '(%.3e contributions/s)' % (sc.Nx, sc.Ny, sc.az_steps, sc.num_ch, sc.pol,
                                         timings['pulse_loop'], w / timings['pulse_loop']))
    return timings


def calibrate_and_add_noise(sc, raw):
    """cal_factor and the complex noise of sar_raw_nompi.py:558-574 on the
    trimmed raw cube [pol][ch][az][rg_orig] (complex128); the noise comes from
    the scene's seeded noise stream."""
    raw = raw * sc.cal_factor
    namp = db2lin(sc.nesz, amplitude=True) / np.sqrt(2.)
    rng = sc.rng_noise
    for ip in range(raw.shape[0]):
        raw[ip] = raw[ip] + namp * (rng.standard_normal(raw[ip].shape) +
                                    1j * rng.standard_normal(raw[ip].shape))
    return raw


def write_raw_file(sc, output_file, raw, nrcs):
    """The RawFile of sar_raw_nompi.py:624-639: raw_data [pol][ch][az][rg_orig],
    NRCS_avg [pol][az] and the scene attributes sar_focus reads back."""
    with raw_io.RawFile(output_file, 'w', raw.shape) as rf:
        rf.set('inc_angle', np.rad2deg(sc.inc_angle))
        rf.set('f0', sc.f0)
        rf.set('num_ch', sc.num_ch)
        rf.set('ant_L', sc.ant_l_tx)
        rf.set('prf', sc.prf)
        rf.set('v_ground', sc.v_ground)
        rf.set('orbit_alt', sc.alt)
        rf.set('sr0', sc.sr_near)
        rf.set('rg_sampling', sc.rg_bw * sc.over_fs)
        rf.set('rg_bw', sc.rg_bw)
        rf.set('raw_data*', raw)
        rf.set('NRCS_avg', nrcs)
        rf.set('b_ati', sc.b_ati)
        rf.set('b_xti', sc.b_xti)


def sarsim(cfg_file, device=0):
    """Raw stage of oceansar_sarsim.sarsim (oceansar_sarsim.py:29-83)."""
    cfg = ConfigFile(cfg_file)
    sim = cfg.sim
    path = getattr(sim, 'path', os.path.dirname(os.path.abspath(cfg_file)))
    if not os.path.isdir(path):
        os.makedirs(path)
    out = None
    if getattr(sim, 'raw_run', True):
        out = sar_raw(cfg.cfg_file_name, os.path.join(path, sim.raw_file),
                      os.path.join(path, sim.ocean_file), getattr(sim, 'ocean_reuse', False),
                      os.path.join(path, getattr(sim, 'errors_file', 'errors.nc')),
                      getattr(sim, 'errors_reuse', False), device=device)
    if getattr(sim, 'proc_run', False):
        from .sar_processor import sar_focus
        sar_focus(cfg.cfg_file_name, os.path.join(path, sim.raw_file),
                  os.path.join(path, sim.proc_file), device=device)
    if getattr(sim, 'insar_run', False):
        from .insar_processor import insar_process
        insar_process(cfg.cfg_file_name, os.path.join(path, sim.proc_file),
                      os.path.join(path, sim.ocean_file), os.path.join(path, sim.insar_file),
                      device=device)
    if getattr(sim, 'ati_run', False):
        from .ati_processor import ati_process
        ati_process(cfg.cfg_file_name, os.path.join(path, sim.insar_file),
                    os.path.join(path, sim.ocean_file), os.path.join(path, sim.ati_file),
                    device=device)
    if getattr(sim, 'L2_wavespectrum_run', False):
        print('sarsim: L2_wavespectrum_run is a downstream stage not built here; skipped')
    return out


if __name__ == '__main__':
    import sys
    if len(sys.argv) < 2:
        raise SystemExit('usage: python -m oceansar_amd.sar_raw <cfg file>')
    sarsim(sys.argv[1])
