"""Restatement of OceanSurface.init's coefficient build (oceansar/surfaces/
ocean.py:131-216) with the counter-based draws of osar_surface_init.

Test infrastructure only. The deterministic part follows the reference's numpy
dtypes (float32 tables, float32 products): |k|, the good-k mask, 1/|k|,
theta, theta_swell, wave_dirspec = kinv S G, swell_dirspec = kinv S_swell and
coefs = Nx Ny sqrt(2 dirspec kx_res ky_res) cg (+ the swell term). The
spectrum, spreading and swell leaves are oceansar_amd.spec.elfouhaily,
oceansar_amd.spread and oceansar_amd.swell_spec, which the golden vectors of
the reference's own functions pin.

The draws replace numpy's unseeded global RNG (ocean.py:201-202, 215-216):
Philox4x32-10 (J. Salmon, M. Moraes, R. Dror, D. Shaw, "Parallel random
numbers: as easy as 1, 2, 3", SC'11; pinned by the Random123 known-answer
vectors), key (seed mod 2^32, seed >> 32), counter (e mod 2^32, e >> 32, 0, 0)
at point e = j Nx + i; cg = sqrt(-2 ln u0) exp(2 pi i u1) / sqrt(2) with
u = (x + 0.5) 2^-32 from words (x0, x1) for the wind sea and (x2, x3) for the
swell, in FP64, rounded to complex64.
"""
import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Ten Philox4x32 rounds on uint32 counter words (arrays) with key (k0, k1)."""
    c = [np.asarray(v, dtype=np.uint64) & _MASK for v in (c0, c1, c2, c3)]
    k0, k1 = np.uint64(k0) & _MASK, np.uint64(k1) & _MASK
    s32 = np.uint64(32)
    for _ in range(10):
        p0 = _M0 * c[0]
        p1 = _M1 * c[2]
        c = [(p1 >> s32) ^ c[1] ^ k0, p1 & _MASK, (p0 >> s32) ^ c[3] ^ k1, p0 & _MASK]
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return [v.astype(np.uint32) for v in c]


def _cg(a, b):
    u0 = (a.astype(np.float64) + 0.5) / 4294967296.
    u1 = (b.astype(np.float64) + 0.5) / 4294967296.
    r = np.sqrt(-2. * np.log(u0)) * (1. / np.sqrt(2.))
    return (r * np.cos(2. * np.pi * u1) + 1j * r * np.sin(2. * np.pi * u1)).astype(np.complex64)


def draws(seed, Ny, Nx):
    """complex64 (wind sea, swell) standard complex Gaussians [Ny, Nx]."""
    e = np.arange(Ny * Nx, dtype=np.uint64)
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    x = philox4x32_10(e & _MASK, e >> np.uint64(32), 0, 0, seed & 0xFFFFFFFF, seed >> 32)
    return _cg(x[0], x[1]).reshape(Ny, Nx), _cg(x[2], x[3]).reshape(Ny, Nx)


def coefficients(surf, seed, spec_model='elfouhaily', spread_model='elfouhaily',
                 swell=None):
    """(wave_coefs complex64, wave_dirspec float32) for the grid and wind of
    `surf` (a SurfaceInit). swell: dict(dir_swell_dir, freq_r, sigf, sigs, Hs)
    for the directional swell term, or None."""
    from oceansar_amd import spec as spec_models
    from oceansar_amd import spread as spread_models
    from oceansar_amd.swell_spec import ardhuin_swell_spec
    kx, ky = np.meshgrid(surf.kx_s, surf.ky_s)
    k = np.sqrt(kx ** 2 + ky ** 2)
    good = np.where(k > np.min(np.array([surf.kx_res, surf.ky_res])) / 2.0)
    kinv = np.zeros(k.shape, dtype=np.float32)
    kinv[good] = 1. / k[good]
    wde = np.float32(surf.wind_dir_eff)
    theta = np.angle(np.exp(1j * (np.arctan2(ky, kx) - wde))).astype(np.float32)
    U = float(surf.wind_U_eff)   # scipy.linalg.norm returns a Python float (ocean.py:120)
    S = np.zeros(k.shape, dtype=np.float32)
    S[good] = spec_models.models[spec_model](k[good], U, surf.wind_fetch)
    G = np.zeros(k.shape, dtype=np.float32)
    G[good] = spread_models.models[spread_model](k[good], theta[good], U, surf.wind_fetch)
    dirspec = kinv * S * G
    cg, cg_s = draws(seed, surf.Ny, surf.Nx)
    n = surf.Nx * surf.Ny
    coefs = (n * np.sqrt(2. * dirspec * surf.kx_res * surf.ky_res) * cg).astype(np.complex64)
    if swell is not None:
        th_s = np.angle(np.exp(1j * np.arctan2(ky, kx))).astype(np.float32)
        sd = np.zeros(k.shape, dtype=np.float32)
        sd[good] = ardhuin_swell_spec(k[good], th_s[good], swell['dir_swell_dir'],
                                      swell['freq_r'], swell['sigf'], swell['sigs'], swell['Hs'])
        sd = kinv * sd
        coefs = coefs + (n * np.sqrt(2. * sd * surf.kx_res * surf.ky_res) *
                         cg_s).astype(np.complex64)
    return coefs, dirspec
