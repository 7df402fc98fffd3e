"""Per-scene host setup of the raw-data generator.

Restates the set-up half of sar_raw (oceansar/radarsim/sar_raw_nompi.py:68-342):
configuration, surface initialisation, FP64 geometry, sizes, closure series,
Bragg spectrum table and range-response table. Everything here runs once per
scene on the host; the per-pulse work runs on the GPU (sar_raw.py).

Decisions where the reference is ambiguous (recorded once here):
  * Cadence is integer. The surface is synthesised at pulses m*K
    ([ocean] surface_update_every, default 1 = every pulse as in
    sar_raw_nompi.py:357-358) and linearly interpolated in between with weight
    (p mod K)/K (the balancer, oceansar/surfaces/balancer.py:131-297).
    Bragg RCS is refreshed at snapshot pulses: every [ocean] rcs_update_every
    pulses if given, else at the first snapshot pulse with
    t_now - t_last > ocean.dt, evaluated in FP64 exactly as
    sar_raw_nompi.py:432 does.
  * The scatterer time index is the integer pulse index (not int(t*Fs)).
  * [sar] n_pulses, when present, overrides the pulse count derived from
    t_span (sar_raw_nompi.py:267-268); t_span itself is unchanged.
  * [ocean] frozen_ocean: the reference skips the surface time update
    (sar_raw_nompi.py:357-358) and so reads fields it never computed
    (np.empty, ocean.py:383-403); here a frozen surface is the surface at t = 0
    for every pulse. The Bragg phase and the scatterer series still advance.
  * Random draws come from numpy.random.SeedSequence([sim] seed).spawn(5) in
    the order: wave coefficients, swell, scatterers minus, scatterers plus,
    noise.
"""
from dataclasses import dataclass

import numpy as np

from . import constants as const
from . import geometry as geosar
from . import spec as spec_models
from .antenna import sinc_bp
from .closure import RandomScatterers, grid_coherence
from .range_profile import N_SINC_SAMPLES, SINC_OVS, calc_sinc_vec
from .spread import SPREAD_IDS
from .surface_init import init_surface, init_surface_device
from .utils import optimize_fftsize


def _get(section, name, default=None):
    return getattr(section, name, default) if section is not None else default


@dataclass
class BraggModel:
    """Bragg (Romeiser 97) model constants (romeiser97.py:38-65, sar_raw_nompi.py:317-338)."""
    k0: float
    d: float
    wind_U: float
    wind_dir: float
    wind_fetch: float
    spread_id: int
    spec_k: np.ndarray        # float64 table abscissa, arange(1, 2 k0, 0.1)
    spec_S: np.ndarray        # float64 spectrum values on spec_k


class SarScene:
    """All host-side inputs of one raw-data simulation."""

    def __init__(self, cfg, seed=None, n_pulses=None, surface=None, device_init=None):
        sar, ocean, srg = cfg.sar, cfg.ocean, cfg.srg
        sim = getattr(cfg, 'sim', None)
        if seed is None:
            seed = int(_get(sim, 'seed', 0))
        self.seed = seed
        streams = np.random.SeedSequence(seed).spawn(5)
        rng = [np.random.Generator(np.random.PCG64(s)) for s in streams]
        self.rng_noise = rng[4]

        # -- [srg] (sar_raw_nompi.py:71-81)
        self.wh_tol = float(srg.wh_tol)
        self.nesz = float(srg.nesz)
        self.use_hmtf = bool(srg.use_hmtf)
        self.spec_enable = bool(srg.scat_spec_enable)
        self.spec_mode = str(_get(srg, 'scat_spec_mode', 'spa'))
        self.bragg_enable = bool(srg.scat_bragg_enable)
        self.bragg_model_name = str(_get(srg, 'scat_bragg_model', 'romeiser97'))
        if self.spec_enable and self.spec_mode != 'spa':
            raise NotImplementedError('RCS mode %s for specular scattering not implemented'
                                      % self.spec_mode)
        if self.bragg_enable and self.bragg_model_name != 'romeiser97':
            raise NotImplementedError('RCS model %s for Bragg scattering not implemented'
                                      % self.bragg_model_name)

        # -- [sar] (sar_raw_nompi.py:84-112)
        self.inc_angle = np.deg2rad(sar.inc_angle)
        self.f0 = float(sar.f0)
        self.pol = str(sar.pol)
        self.pols = {'DP': ['hh', 'vv'], 'hh': ['hh']}.get(self.pol, ['vv'])
        self.prf = float(sar.prf)
        self.num_ch = int(sar.num_ch)
        self.alt = float(sar.alt)
        v_ground = sar.v_ground
        self.rg_bw = float(sar.rg_bw)
        self.over_fs = float(sar.over_fs)
        self.heading = float(_get(sar, 'heading', 0.))

        # -- surface (sar_raw_nompi.py:126-190, ocean.py:29-242)
        depth = _get(ocean, 'depth', None)
        if isinstance(depth, str):
            depth = None if depth.lower() == 'none' else float(depth)
        swell_dir_enable = bool(_get(ocean, 'swell_dir_enable', False))
        swell_enable = bool(_get(ocean, 'swell_enable', False))
        if device_init is not None and surface is None:
            # spectrum and coefficients built in HBM (osar_surface_init), keyed by
            # the wind-sea stream of the scene seed; device_init is the device index
            surface = init_surface_device(
                float(ocean.Lx), float(ocean.Ly), float(ocean.dx), float(ocean.dy),
                str(ocean.spec_model), str(ocean.spread_model),
                np.deg2rad(ocean.wind_dir), float(ocean.wind_fetch), float(ocean.wind_U),
                float(ocean.current_mag), np.deg2rad(ocean.current_dir),
                seed=int(streams[0].generate_state(1, np.uint64)[0]),
                dir_swell_dir=_get(ocean, 'dir_swell_dir', 0),
                freq_r=_get(ocean, 'freq_r', 100), sigf=_get(ocean, 'sigf', 1),
                sigs=_get(ocean, 'sigs', 1), Hs=_get(ocean, 'Hs', 2),
                swell_dir_enable=swell_dir_enable, swell_enable=swell_enable,
                opt_res=bool(_get(ocean, 'opt_res', True)),
                fft_max_prime=int(_get(ocean, 'fft_max_prime', 2)),
                choppy_enable=bool(_get(ocean, 'choppy_enable', False)),
                depth=depth, cutoff_wl=_get(ocean, 'cutoff_wl', 'auto'), device=int(device_init))
        self.surface = surface if surface is not None else init_surface(
            float(ocean.Lx), float(ocean.Ly), float(ocean.dx), float(ocean.dy),
            str(ocean.spec_model), str(ocean.spread_model),
            np.deg2rad(ocean.wind_dir), float(ocean.wind_fetch), float(ocean.wind_U),
            float(ocean.current_mag), np.deg2rad(ocean.current_dir),
            rng_wave=rng[0], rng_swell=rng[1],
            dir_swell_dir=_get(ocean, 'dir_swell_dir', 0),
            freq_r=_get(ocean, 'freq_r', 100), sigf=_get(ocean, 'sigf', 1),
            sigs=_get(ocean, 'sigs', 1), Hs=_get(ocean, 'Hs', 2),
            swell_dir_enable=swell_dir_enable, swell_enable=swell_enable,
            swell_ampl=_get(ocean, 'swell_ampl', 0.),
            swell_dir=np.deg2rad(_get(ocean, 'swell_dir', 0.)),
            swell_wl=_get(ocean, 'swell_wl', 1.),
            opt_res=bool(_get(ocean, 'opt_res', True)),
            fft_max_prime=int(_get(ocean, 'fft_max_prime', 2)),
            choppy_enable=bool(_get(ocean, 'choppy_enable', False)),
            depth=depth, cutoff_wl=_get(ocean, 'cutoff_wl', 'auto'))
        surf = self.surface
        self.Nx, self.Ny = surf.Nx, surf.Ny

        # -- geometry (sar_raw_nompi.py:231-246)
        self.sr0 = geosar.inc_to_sr(self.inc_angle, self.alt)
        self.gr0 = geosar.inc_to_gr(self.inc_angle, self.alt)
        gr = surf.x + self.gr0
        sr, inc, _ = geosar.gr_to_geo(gr, self.alt)
        self.rg_sampling = self.rg_bw * self.over_fs
        self.sr_near = sr[0] - self.wh_tol + const.c / 2 / self.rg_sampling
        sr = sr - np.min(sr)
        self.sr = np.ascontiguousarray(sr, dtype=np.float64)          # [Nx]
        self.inc = np.ascontiguousarray(inc, dtype=np.float64)        # [Nx]
        self.sin_inc = np.sin(self.inc)
        self.cos_inc = np.cos(self.inc)
        self.y64 = surf.y.astype(np.float64)                          # [Ny]

        # -- wavelength, resolution, baselines (sar_raw_nompi.py:249-262)
        self.l0 = const.c / self.f0
        self.k0 = 2. * np.pi * self.f0 / const.c
        self.sr_res = const.c / (2. * self.rg_bw)
        self.ant_l_tx = float(sar.ant_L_tx)
        self.ant_l_rx = float(sar.ant_L_rx)
        b_ati = sar.b_ati
        if not isinstance(b_ati, np.ndarray):
            b_ati = np.arange(self.num_ch) * b_ati
        b_xti = sar.b_xti
        if not isinstance(b_xti, np.ndarray):
            b_xti = np.arange(self.num_ch) * b_xti
        self.b_ati = np.asarray(b_ati, dtype=np.float64)[:self.num_ch]
        self.b_xti = np.asarray(b_xti, dtype=np.float64)[:self.num_ch]
        if self.b_ati.size != self.num_ch or self.b_xti.size != self.num_ch:
            raise ValueError('b_ati/b_xti need num_ch entries')

        # -- timing and sizes (sar_raw_nompi.py:264-278)
        if isinstance(v_ground, str) and v_ground == 'auto':
            v_ground = geosar.orbit_to_vel(self.alt, ground=True)
        self.v_ground = float(v_ground)
        self.t_step = 1. / self.prf
        self.t_span = (1.5 * (self.sr0 * self.l0 / self.ant_l_tx) + surf.Ly) / self.v_ground
        az_steps = int(np.floor(self.t_span / self.t_step))
        override = n_pulses if n_pulses is not None else _get(sar, 'n_pulses', None)
        self.az_steps = int(override) if override is not None else az_steps
        self.az_span = (self.t_span * self.v_ground + surf.Ly) / self.gr0
        max_sr = (np.max(self.sr) + self.wh_tol +
                  (np.max(surf.y) + (self.t_span / 2.) * self.v_ground) ** 2. / (2. * self.sr0))
        self.min_sr = np.min(self.sr) - self.wh_tol
        self.rg_samp_orig = int(np.ceil(((max_sr - self.min_sr) / self.sr_res) * self.over_fs))
        self.rg_samp = int(optimize_fftsize(self.rg_samp_orig))
        self.binsize = self.sr_res / self.over_fs
        self.surface_area = float(surf.dx * surf.dy * surf.Nx * surf.Ny)

        # -- cadence (integer; see module docstring)
        K = int(_get(ocean, 'surface_update_every', 1))
        if K < 1:
            raise ValueError('surface_update_every must be >= 1')
        self.K = K
        self.ocean_dt = float(ocean.dt)
        self.frozen = bool(_get(ocean, 'frozen_ocean', False))
        k_rcs = _get(ocean, 'rcs_update_every', None)
        self.rcs_refresh = self._rcs_schedule(None if k_rcs is None else int(k_rcs))
        # rcs_slot[p] = index of the RCS refresh in force at pulse p
        slot = np.zeros(self.az_steps, dtype=np.int32)
        slot[self.rcs_refresh[1:]] = 1
        self.rcs_slot = np.cumsum(slot).astype(np.int32)

        # -- range response (sar_raw_nompi.py:119-121)
        self.n_sinc = N_SINC_SAMPLES
        self.sinc_ovs = SINC_OVS
        self.sinc_vec = calc_sinc_vec(self.n_sinc, self.sinc_ovs, Fs=self.over_fs)

        # -- Bragg closure and model (sar_raw_nompi.py:305-338)
        self.bragg = None
        if self.bragg_enable:
            self.tau_c = grid_coherence(float(ocean.wind_U), float(surf.dx), self.f0)
            self.scat_m = RandomScatterers(self.tau_c, (self.Ny, self.Nx), self.prf, rng[2])
            self.scat_p = RandomScatterers(self.tau_c, (self.Ny, self.Nx), self.prf, rng[3])
            k_b = 2. * self.k0 * self.sin_inc
            self.c_b = self.sin_inc * np.sqrt(const.g / k_b + 0.072e-3 * k_b)
            self.cap_rate = (2 * self.k0) * self.t_step * self.c_b      # [Nx], times (p+1)
            current_dir = np.deg2rad(ocean.current_dir)
            current_vec = float(ocean.current_mag) * np.array([np.cos(current_dir),
                                                               np.sin(current_dir)])
            U_dir = np.deg2rad(ocean.wind_dir)
            U_vec = float(ocean.wind_U) * np.array([np.cos(U_dir), np.sin(U_dir)])
            U_eff_vec = U_vec - current_vec
            spec_name = str(srg.scat_bragg_spec)
            spread_name = str(srg.scat_bragg_spread)
            if spread_name not in SPREAD_IDS:
                raise NotImplementedError('Bragg spreading %s not supported' % spread_name)
            wind_U = float(np.linalg.norm(U_eff_vec))
            spec_k = np.arange(1, 2 * self.k0, 0.1)
            spec_S = np.asarray(spec_models.models[spec_name](spec_k, wind_U, surf.wind_fetch),
                                dtype=np.float64)
            self.bragg = BraggModel(k0=self.k0, d=float(srg.scat_bragg_d), wind_U=wind_U,
                                    wind_dir=float(np.arctan2(U_eff_vec[1], U_eff_vec[0])),
                                    wind_fetch=float(surf.wind_fetch),
                                    spread_id=SPREAD_IDS[spread_name],
                                    spec_k=spec_k, spec_S=spec_S)

        # -- calibration (sar_raw_nompi.py:557-563)
        az_axis = np.arange(-self.t_span / 2. * self.v_ground, self.t_span / 2. * self.v_ground,
                            self.sr0 * const.c / (np.pi * self.f0 * self.ant_l_tx * 10.))
        pattern = (sinc_bp(az_axis / self.sr0, self.ant_l_tx, self.f0, field=True) *
                   sinc_bp(az_axis / self.sr0, self.ant_l_rx, self.f0, field=True))
        self.cal_factor = (1. / np.sqrt(np.trapezoid(np.abs(pattern) ** 2., az_axis) *
                                        self.sr_res / np.sin(self.inc_angle)))

    # ------------------------------------------------------------------
    def _rcs_schedule(self, k_rcs):
        """Pulses at which the Bragg RCS is refreshed."""
        snaps = np.arange(0, self.az_steps, self.K)
        if k_rcs is not None:
            if k_rcs % self.K:
                raise ValueError('rcs_update_every must be a multiple of surface_update_every')
            return np.arange(0, self.az_steps, k_rcs, dtype=np.int64)
        out = []
        t_last = -1.
        for p in snaps:
            t_now = p * self.t_step
            if (t_now - t_last) > self.ocean_dt:
                out.append(int(p))
                t_last = t_now
        return np.asarray(out, dtype=np.int64)

    # ------------------------------------------------------------------
    @property
    def n_snapshots(self):
        """Surface snapshots needed: pulses 0..az_steps-1 use m and m+1."""
        return (self.az_steps - 1) // self.K + 2 if self.K > 1 else self.az_steps

    def snapshot_time(self, m):
        """Time of surface snapshot m: m K t_step, or 0 for a frozen ocean."""
        if self.frozen:
            return 0.
        return float(m * self.K) * self.t_step

    def az_now(self, p):
        """Platform along-track position at pulse p (sar_raw_nompi.py:352-353)."""
        return (p * self.t_step - self.t_span / 2.) * self.v_ground

    def az_proj_row(self, p):
        az = self.y64 - self.az_now(p)
        return np.arcsin(az / self.gr0)

    @property
    def contributions(self):
        """Facet-pulse contributions W = facets * pulses * channels * pols."""
        return self.Nx * self.Ny * self.az_steps * self.num_ch * len(self.pols)
