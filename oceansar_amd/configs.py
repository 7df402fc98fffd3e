"""Synthetic scene configurations in the reference's ``.cfg`` format.

Section and key names follow par/sim_nompi_luttest.cfg (the only shipped cfg
with every key sar_raw_nompi.py reads). Three optional keys are added:
[sim] seed, [sar] n_pulses, [ocean] surface_update_every / rcs_update_every.
The antenna length is chosen so that the reference's own pulse-count formula
(sar_raw_nompi.py:267-268) gives the requested number of pulses.
"""
import numpy as np

from . import constants as const
from . import geometry as geosar

_TEMPLATE = """[sim]
seed={seed}
raw_run=True
raw_file=raw_data.nc
ocean_file=ocean.nc
ocean_reuse=False
errors_file=errors.nc
errors_reuse=False
proc_file=proc_data.nc
insar_file=insar_l1b.nc
ati_file=ati_data.txt
proc_run={proc_run}
insar_run={insar_run}
ati_run={ati_run}
L2_wavespectrum_run=False

[sar]
f0={f0}
inc_angle={inc}
pol={pol}
prf={prf}
rg_bw={rg_bw}
over_fs=1.1
squint=0
alt={alt}
v_ground={v}
heading=0
ant_L_tx={antL}
ant_L_rx={antL}
num_ch={nch}
b_ati={b_ati}
b_xti={b_xti}
n_pulses={npulses}
use_errors=False
sigma_n_tx=0
phase_n_tx=0
sigma_beta_tx=0
phase_beta_tx=0
sigma_n_rx=0
phase_n_rx=0
sigma_beta_rx=0
phase_beta_rx=0

[ocean]
Lx={L}
Ly={L}
dx={dx}
dy={dx}
opt_res=True
fft_max_prime=2
dt={dt}
surface_update_every={K}
depth=1000
wind_dir={wdir}
wind_U={U}
wind_fetch=500e3
swell_dir_enable={swell}
sigf=0.007
freq_r={freq_r}
dir_swell_dir={swell_dir}
sigs=8
Hs={Hs}
swell_enable=False
swell_wl=80
swell_dir=45
swell_ampl=0.5
spec_model=elfouhaily
spread_model={spread}
cutoff_wl=auto
current_mag=0
current_dir=90.
choppy_enable={choppy}
frozen_ocean={frozen}
use_buoy_data=False

[processing]
az_weighting=1
doppler_bw=1200.
plot_format=png
plot_tex=False
plot_save=False
plot_path=processing_plots/
plot_raw=False
plot_rcmc_dopp=False
plot_rcmc_time=False
plot_image_valid=False

[insar]
rg_ml={rg_ml}
az_ml={az_ml}
ml_win={ml_win}
flatten={flatten}
plot_save=False
plot_path=insar_plots/
plot_format=png
plot_tex=False
plot_surface=False
plot_proc_ampl=False
plot_coh=False
plot_coh_all=False

[ati]
rg_ml=8
az_ml=2
ml_win={ml_win}
plot_save=False
plot_path=ati_plots/
plot_format=png
plot_tex=False
plot_surface=False
plot_proc_ampl=False
plot_coh=False
plot_coh_all=False
plot_ati_phase=False
plot_ati_phase_all=False
plot_vel_hist=False
plot_vel=False

[srg]
use_hmtf={hmtf}
nesz=-30.
wh_tol=2.73
scat_spec_enable={specular}
scat_spec_mode=spa
scat_bragg_enable=True
scat_bragg_model=romeiser97
scat_bragg_spec=elfouhaily
scat_bragg_spread=elfouhaily
scat_bragg_d=0.125
use_lut=False
"""


def antenna_length_for(n_pulses, prf, v, Ly, inc_deg, alt, f0):
    """Antenna length giving floor(t_span * prf) == n_pulses (half-pulse margin)."""
    sr0 = geosar.inc_to_sr(np.deg2rad(inc_deg), alt)
    l0 = const.c / f0
    t_span = (n_pulses + 0.5) / prf
    return 1.5 * sr0 * l0 / (t_span * v - Ly)


def _arr(vals):
    # A one-element list like [0] would parse as a section header (cfg.py:90),
    # so a single baseline is written as a scalar, which sar_raw expands with
    # arange(num_ch) * value (sar_raw_nompi.py:256-262).
    if len(vals) == 1:
        return '%g' % vals[0]
    return '[' + ', '.join('%g' % v for v in vals) + ']'


def cfg_text(L, dx, n_pulses, K, U, wind_dir=0., num_ch=1, b_ati=None, pol='vv',
             spread='elfouhaily', specular=False, swell=False, seed=1,
             f0=5.405e9, inc=35., prf=1800., rg_bw=100e6, alt=514e3, v=7600.,
             swell_wl=200., Hs=2., swell_dir=0., choppy=False, dt=None, b_xti=None,
             frozen=False, hmtf=False, stages=(), rg_ml=5, az_ml=5, ml_win='hanning',
             flatten=True):
    antL = antenna_length_for(n_pulses, prf, v, L, inc, alt, f0)
    b_ati = [0.] * num_ch if b_ati is None else list(b_ati)[:num_ch]
    freq_r = np.sqrt(const.g * 2 * np.pi / swell_wl) / (2 * np.pi)
    if dt is None:
        dt = K / prf * 0.999
    return _TEMPLATE.format(
        seed=seed, f0=f0, inc=inc, pol=pol, prf=prf, rg_bw=rg_bw, alt=alt, v=v,
        antL=repr(float(antL)), nch=num_ch, b_ati=_arr(b_ati),
        b_xti=_arr([0.] * num_ch if b_xti is None else list(b_xti)[:num_ch]),
        npulses=n_pulses, L=L, dx=dx, dt=repr(float(dt)), K=K, wdir=wind_dir, U=U,
        swell=swell, freq_r=repr(float(freq_r)), swell_dir=swell_dir, Hs=Hs,
        spread=spread, choppy=choppy, specular=specular, frozen=frozen, hmtf=hmtf,
        proc_run='proc' in stages, insar_run='insar' in stages, ati_run='ati' in stages,
        rg_ml=rg_ml, az_ml=az_ml, ml_win=ml_win, flatten=flatten)


WORKLOADS = ('small', 'ati_swell', 'large')


def workload(name, **over):
    """cfg text for the measured workloads; keyword overrides go to cfg_text.

    small:     256^2 facets at 1 m, 1 channel, 1024 pulses, K=16, U10=10, cos^2 spread.
    ati_swell: 1024^2 at 0.5 m, 2-channel ATI (0/10 m), 4096 pulses, K=8, U10=7 at
               45 deg, 200 m directional swell (Hs 2 m), Bragg + specular.
    large:     2048^2 at 0.5 m, 4 channels (0/5/10/20 m), 8192 pulses, K=16, U10=12.
    """
    index = {'small': 1, 'ati_swell': 2, 'large': 3}.get(name)
    if index == 1:
        kw = dict(L=256., dx=1., n_pulses=1024, K=16, U=10., wind_dir=0., num_ch=1,
                  spread='cos2s', seed=1)
    elif index == 2:
        kw = dict(L=512., dx=0.5, n_pulses=4096, K=8, U=7., wind_dir=45., num_ch=2,
                  b_ati=[0., 10.], specular=True, swell=True, swell_wl=200., Hs=2.,
                  swell_dir=45., seed=2)
    elif index == 3:
        kw = dict(L=1024., dx=0.5, n_pulses=8192, K=16, U=12., wind_dir=0., num_ch=4,
                  b_ati=[0., 5., 10., 20.], seed=3)
    else:
        raise ValueError('unknown workload %r (known: %s)' % (name, ', '.join(WORKLOADS)))
    kw.update(over)
    return cfg_text(**kw)


def write(path, text):
    with open(path, 'w') as fh:
        fh.write(text)
    return path


def load_scene(text, tmpdir, name='scene.cfg', **kw):
    """Write cfg text, parse it with ConfigFile and build a SarScene."""
    import os
    from .cfg import ConfigFile
    from .scene import SarScene
    p = write(os.path.join(str(tmpdir), name), text)
    return SarScene(ConfigFile(p), **kw)
