"""Host-side ocean surface initialisation (wave spectrum -> random coefficients).

Restates OceanSurface.init (oceansar/surfaces/ocean.py:29-242) with the same
numpy dtypes (float32 grids and spectra, complex64 coefficients) so that the
coefficients are what the reference would build from the same random draws.
The random draws are explicit inputs (numpy Generators) instead of the
reference's unseeded global RNG (ocean.py:201-202, 215-216, 229).
"""
from dataclasses import dataclass, field

import numpy as np
from scipy import linalg

from . import constants as const
from . import spec as spec_models
from . import spread as spread_models
from .swell_spec import ardhuin_swell_spec
from .utils import optimize_fftsize


@dataclass
class SurfaceInit:
    Lx: float
    Ly: float
    Nx: int
    Ny: int
    dx: float
    dy: float
    x: np.ndarray            # float32 [Nx]  (linspace(-Lx/2, Lx/2, Nx), ocean.py:113)
    y: np.ndarray            # float32 [Ny]
    kx_s: np.ndarray         # float32 [Nx]  (2*pi*fftfreq, ocean.py:131)
    ky_s: np.ndarray         # float32 [Ny]
    kx_res: np.float32
    ky_res: np.float32
    wave_coefs: np.ndarray   # complex64 [Ny, Nx]
    wave_dirspec: np.ndarray # float32 [Ny, Nx]
    rmss_x: float
    rmss_y: float
    wind_U: float
    wind_dir: float
    wind_U_eff: float
    wind_dir_eff: float
    wind_fetch: float
    current: np.ndarray      # float32 [2]
    depth: object            # None = deep water
    choppy: bool
    swell_enable: bool = False
    swell_ampl: float = 0.
    swell_dir: float = 0.
    swell_wl: float = 0.
    swell_k: float = 0.
    swell_kx: float = 0.
    swell_ky: float = 0.
    swell_omega: float = 0.
    swell_ph0: float = 0.
    extra: dict = field(default_factory=dict)


def _complex_gaussian(rng, shape):
    # (N(0,1) + i N(0,1)) / sqrt(2) as complex64 (ocean.py:201-202)
    re = rng.standard_normal(size=shape)
    im = rng.standard_normal(size=shape)
    return (1. / np.sqrt(2) * (re + 1j * im)).astype(np.complex64)


def _grid(Lx, Ly, dx, dy, wind_dir, wind_U, current_mag, current_dir, opt_res, fft_max_prime,
          cutoff_wl):
    """Sizes, grids, current and effective wind of ocean.py:100-137."""
    if cutoff_wl and cutoff_wl != 'auto':
        # ocean.py:185-198 slices with float indices and fails for any value
        raise NotImplementedError("only cutoff_wl='auto' is supported")
    Nx = int(Lx / dx)
    Ny = int(Ly / dy)
    if opt_res:
        Nx = int(optimize_fftsize(Nx, fft_max_prime))
        Ny = int(optimize_fftsize(Ny, fft_max_prime))
        dx = np.float32(Lx / float(Nx))
        dy = np.float32(Ly / float(Ny))
    x = np.linspace(-Lx / 2., Lx / 2., Nx, dtype=np.float32)
    y = np.linspace(-Ly / 2., Ly / 2., Ny, dtype=np.float32)
    current = (current_mag * np.array([np.cos(current_dir),
                                       np.sin(current_dir)])).astype(np.float32)
    U_eff_vec = (wind_U * np.array([np.cos(wind_dir), np.sin(wind_dir)]) -
                 current).astype(np.float32)
    kx_s = (2 * np.pi * np.fft.fftfreq(Nx, dx)).astype(np.float32)
    ky_s = (2 * np.pi * np.fft.fftfreq(Ny, dy)).astype(np.float32)
    return dict(Nx=Nx, Ny=Ny, dx=dx, dy=dy, x=x, y=y, current=current,
                wind_U_eff=linalg.norm(U_eff_vec),
                wind_dir_eff=np.arctan2(U_eff_vec[1], U_eff_vec[0]),
                kx_s=kx_s, ky_s=ky_s,
                kx_res=kx_s[1] - kx_s[0] if Nx > 1 else np.float32(0),
                ky_res=ky_s[1] - ky_s[0] if Ny > 1 else np.float32(0))


def init_surface(Lx, Ly, dx, dy, spec_model, spread_model, wind_dir, wind_fetch,
                 wind_U, current_mag, current_dir, rng_wave, rng_swell=None,
                 dir_swell_dir=0, freq_r=100, sigf=1, sigs=1, Hs=2,
                 swell_dir_enable=False, swell_enable=False, swell_ampl=0.,
                 swell_dir=0., swell_wl=0., opt_res=True, fft_max_prime=2,
                 choppy_enable=False, depth=None, cutoff_wl='auto'):
    """Build the surface grid and complex wave coefficients (ocean.py:100-242)."""
    g = _grid(Lx, Ly, dx, dy, wind_dir, wind_U, current_mag, current_dir, opt_res,
              fft_max_prime, cutoff_wl)
    Nx, Ny, dx, dy, x, y = g['Nx'], g['Ny'], g['dx'], g['dy'], g['x'], g['y']
    current, wind_U_eff, wind_dir_eff = g['current'], g['wind_U_eff'], g['wind_dir_eff']
    kx_s, ky_s, kx_res, ky_res = g['kx_s'], g['ky_s'], g['kx_res'], g['ky_res']
    kx, ky = np.meshgrid(kx_s, ky_s)
    k = np.sqrt(kx ** 2 + ky ** 2)
    good_k = np.where(k > np.min(np.array([kx_res, ky_res])) / 2.0)
    kinv = np.zeros(k.shape, dtype=np.float32)
    kinv[good_k] = 1. / k[good_k]
    theta = np.angle(np.exp(1j * (np.arctan2(ky, kx) - wind_dir_eff))).astype(np.float32)
    theta_swell = np.angle(np.exp(1j * np.arctan2(ky, kx))).astype(np.float32)

    if spec_model not in spec_models.models:
        raise NotImplementedError('%s spectrum function not implemented' % spec_model)
    if spread_model not in spread_models.models or spread_models.models[spread_model] is None:
        raise NotImplementedError('%s spreading function not implemented' % spread_model)
    wave_spec = np.zeros(k.shape, dtype=np.float32)
    wave_spec[good_k] = spec_models.models[spec_model](k[good_k], wind_U_eff, wind_fetch)
    wave_spread = np.zeros(k.shape, dtype=np.float32)
    wave_spread[good_k] = spread_models.models[spread_model](k[good_k], theta[good_k],
                                                             wind_U_eff, wind_fetch)
    wave_dirspec = kinv * wave_spec * wave_spread
    rmss_x = float(np.sqrt(np.sum(wave_dirspec * kx ** 2 * kx_res * ky_res)))
    rmss_y = float(np.sqrt(np.sum(wave_dirspec * ky ** 2 * kx_res * ky_res)))

    random_cg = _complex_gaussian(rng_wave, (Ny, Nx))
    wave_coefs = (Nx * Ny * np.sqrt(2. * wave_dirspec * kx_res * ky_res) *
                  random_cg).astype(np.complex64)
    if swell_dir_enable:
        if rng_swell is None:
            raise ValueError('directional swell needs its own random stream')
        swell_dirspec = np.zeros(k.shape, dtype=np.float32)
        swell_dirspec[good_k] = ardhuin_swell_spec(k[good_k], theta_swell[good_k],
                                                   dir_swell_dir, freq_r, sigf, sigs, Hs)
        swell_dirspec = kinv * swell_dirspec
        random_swell = _complex_gaussian(rng_swell, (Ny, Nx))
        swell_coefs = (Nx * Ny * np.sqrt(2. * swell_dirspec * kx_res * ky_res) *
                       random_swell).astype(np.complex64)
        wave_coefs = wave_coefs + swell_coefs

    s = SurfaceInit(Lx=Lx, Ly=Ly, Nx=Nx, Ny=Ny, dx=dx, dy=dy, x=x, y=y,
                    kx_s=kx_s, ky_s=ky_s, kx_res=kx_res, ky_res=ky_res,
                    wave_coefs=wave_coefs, wave_dirspec=wave_dirspec,
                    rmss_x=rmss_x, rmss_y=rmss_y, wind_U=wind_U, wind_dir=wind_dir,
                    wind_U_eff=float(wind_U_eff), wind_dir_eff=float(wind_dir_eff),
                    wind_fetch=wind_fetch, current=current, depth=depth,
                    choppy=bool(choppy_enable))
    if swell_enable:
        # Monochromatic swell (ocean.py:218-233)
        if rng_swell is None:
            raise ValueError('monochromatic swell needs its own random stream')
        s.swell_enable = True
        s.swell_ampl = float(swell_ampl)
        s.swell_dir = float(swell_dir)
        s.swell_wl = float(swell_wl)
        s.swell_k = float(np.float32(2. * np.pi / swell_wl))
        s.swell_kx = float(np.float32(s.swell_k * np.cos(swell_dir)))
        s.swell_ky = float(np.float32(s.swell_k * np.sin(swell_dir)))
        if depth is None:
            s.swell_omega = float(np.float32(np.sqrt(s.swell_k * const.g)))
        else:
            s.swell_omega = float(np.float32(np.sqrt(s.swell_k * const.g *
                                                     np.tanh(s.swell_k * depth))))
        s.swell_ph0 = float(np.float32(rng_swell.uniform(0., 2. * np.pi)))
    return s


def init_surface_device(Lx, Ly, dx, dy, spec_model, spread_model, wind_dir, wind_fetch,
                        wind_U, current_mag, current_dir, seed, dir_swell_dir=0, freq_r=100,
                        sigf=1, sigs=1, Hs=2, swell_dir_enable=False, swell_enable=False,
                        opt_res=True, fft_max_prime=2, choppy_enable=False, depth=None,
                        cutoff_wl='auto', device=0, ctx=None):
    """init_surface with the spectrum and the coefficients built on the device
    (osar_surface_init): wave_coefs and wave_dirspec of the result are device
    tensors ([Ny, Nx] complex64 and float32) that never pass through the host.
    The draws come from Philox4x32-10 keyed by `seed` (an integer below 2^64),
    so the coefficients are not those init_surface draws with numpy Generators.
    """
    import ctypes
    import torch
    from . import _lib
    if swell_enable:
        raise NotImplementedError('monochromatic swell is not supported on the GPU path')
    if spec_model not in spec_models.models:
        raise NotImplementedError('%s spectrum function not implemented' % spec_model)
    if spread_model not in spread_models.SPREAD_IDS:
        raise NotImplementedError('%s spreading function not implemented' % spread_model)
    g = _grid(Lx, Ly, dx, dy, wind_dir, wind_U, current_mag, current_dir, opt_res,
              fft_max_prime, cutoff_wl)
    Nx, Ny = g['Nx'], g['Ny']
    ctx = ctx if ctx is not None else _lib.context(device)
    dev = torch.device('cuda', ctx.device)
    kx = torch.as_tensor(g['kx_s'], device=dev)
    ky = torch.as_tensor(g['ky_s'], device=dev)
    coefs = torch.empty((Ny, Nx), dtype=torch.complex64, device=dev)
    dirspec = torch.empty((Ny, Nx), dtype=torch.float32, device=dev)
    sums = torch.empty(2, dtype=torch.float64, device=dev)
    d = _lib.InitDesc()
    d.nx, d.ny = Nx, Ny
    d.spread = spread_models.SPREAD_IDS[spread_model]
    d.swell = 1 if swell_dir_enable else 0
    d.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    d.kx_res, d.ky_res = float(g['kx_res']), float(g['ky_res'])
    d.wind_dir_eff = float(g['wind_dir_eff'])
    e = spec_models.elfouhaily_scalars(g['wind_U_eff'], wind_fetch)
    d.k_p, d.c_p, d.omega_w = float(e['k_p']), float(e['c_p']), float(e['Omega'])
    d.alpha_p, d.sigma, d.gamma = float(e['alpha_p']), float(e['sigma']), float(e['gamma'])
    d.alpha_m, d.c_m, d.k_m = float(e['alpha_m']), float(e['c_m']), float(e['k_m'])
    # the host spreading evaluates Delta(k) in FP64 from float(U_eff)
    sp = spread_models.elfouhaily_delta_scalars(float(g['wind_U_eff']), float(wind_fetch))
    d.a_0, d.a_p, d.a_m = float(sp['a_0']), float(sp['a_p']), float(sp['a_m'])
    d.cK_p, d.cK_m = float(sp['cK_p']), float(sp['cK_m'])
    d.swell_dir, d.freq_r = float(np.radians(dir_swell_dir)), float(freq_r)
    d.sigf, d.sigs = float(sigf), float(np.radians(sigs))
    d.swell_amp = float((Hs / 4) ** 2 / sigf / np.sqrt(2 * np.pi))
    d.kx, d.ky = kx.data_ptr(), ky.data_ptr()
    d.coefs, d.dirspec, d.rmss_sums = coefs.data_ptr(), dirspec.data_ptr(), sums.data_ptr()
    _lib.check(ctx.lib.osar_surface_init(ctx.h, ctypes.byref(d), _lib.stream_ptr()))
    sx, sy = sums.cpu().numpy()      # waits for the launch
    return SurfaceInit(Lx=Lx, Ly=Ly, Nx=Nx, Ny=Ny, dx=g['dx'], dy=g['dy'], x=g['x'], y=g['y'],
                       kx_s=g['kx_s'], ky_s=g['ky_s'], kx_res=g['kx_res'], ky_res=g['ky_res'],
                       wave_coefs=coefs, wave_dirspec=dirspec,
                       rmss_x=float(np.sqrt(sx)), rmss_y=float(np.sqrt(sy)), wind_U=wind_U,
                       wind_dir=wind_dir, wind_U_eff=float(g['wind_U_eff']),
                       wind_dir_eff=float(g['wind_dir_eff']), wind_fetch=wind_fetch,
                       current=g['current'], depth=depth, choppy=bool(choppy_enable),
                       extra={'seed': int(seed)})


def synthesis_tables(s):
    """FP64 wave-number tables shared by the GPU synthesis and the oracle.

    k, kx/k, ky/k and 1/k are formed in FP64 from the float32 kx, ky grid;
    the good-k mask is the reference's k > min(dk)/2 (ocean.py:141-148).
    """
    kx = s.kx_s.astype(np.float64)[None, :]
    ky = s.ky_s.astype(np.float64)[:, None]
    k = np.sqrt(kx ** 2 + ky ** 2)
    thr = float(np.min(np.array([s.kx_res, s.ky_res]))) / 2.0
    good = k > thr
    return kx, ky, k, good, thr
