"""Pinning the oracle's surface synthesis (ocean.py:418-526) by closed forms.

A spectrum with a single non-zero coefficient c at wave vector k gives
analytic fields: with A = c exp(-i w t)/(Nx Ny) and phi = kx x + ky y on the
FFT grid, Dz = Re(A e^{i phi}), Dx = -(kx/k) Im(A e^{i phi}),
Diffx = -kx Im(A e^{i phi}), Diffxx = -kx^2 Re(A e^{i phi}), ... and the
choppy slope divides by 1 - (kx^2/k) Re(A e^{i phi}); w = sqrt(g k tanh(k d)).
These fix the multipliers, their signs, the dispersion relation and the sign
of the time phase. Parseval fixes the normalisation over a random spectrum.
"""
import numpy as np
import pytest

import oracle.surface as osurf
from oceansar_amd.surface_init import SurfaceInit

G = 9.80665


def _surf(Nx, Ny, dx, dy, coefs, depth=None, choppy=False, current=(0., 0.)):
    kx_s = (2 * np.pi * np.fft.fftfreq(Nx, dx)).astype(np.float32)
    ky_s = (2 * np.pi * np.fft.fftfreq(Ny, dy)).astype(np.float32)
    return SurfaceInit(Lx=Nx * dx, Ly=Ny * dy, Nx=Nx, Ny=Ny, dx=dx, dy=dy,
                       x=np.zeros(Nx, np.float32), y=np.zeros(Ny, np.float32),
                       kx_s=kx_s, ky_s=ky_s, kx_res=kx_s[1] - kx_s[0], ky_res=ky_s[1] - ky_s[0],
                       wave_coefs=coefs.astype(np.complex64), wave_dirspec=None,
                       rmss_x=0., rmss_y=0., wind_U=0., wind_dir=0., wind_U_eff=0.,
                       wind_dir_eff=0., wind_fetch=0., current=np.array(current, np.float32),
                       depth=depth, choppy=choppy)


@pytest.mark.parametrize('depth', [None, 15.0])
@pytest.mark.parametrize('mode', [(3, 5), (-7, 2), (0, 4), (6, 0)])
def test_single_mode_closed_form(mode, depth):
    Nx, Ny, dx, dy = 64, 32, 1.5, 2.0
    i0, j0 = mode[0] % Nx, mode[1] % Ny
    c = np.zeros((Ny, Nx), complex)
    c[j0, i0] = 3.0 - 1.2j
    s = _surf(Nx, Ny, dx, dy, c, depth=depth)
    kx = float(s.kx_s[i0])
    ky = float(s.ky_s[j0])
    k = np.hypot(kx, ky)
    w = np.sqrt(G * k) if depth is None else np.sqrt(G * k * np.tanh(k * depth))
    t = 2.37
    f = osurf.synth(s, t, ('D', 'Diff', 'Diff2'))
    n = np.arange(Nx)[None, :]
    m = np.arange(Ny)[:, None]
    phi = 2 * np.pi * (i0 * n / Nx + j0 * m / Ny)
    A = complex(np.complex64(c[j0, i0])) * np.exp(-1j * w * t) / (Nx * Ny)
    z = A * np.exp(1j * phi)
    exp = {'Dz': z.real, 'Dx': -(kx / k) * z.imag, 'Dy': -(ky / k) * z.imag,
           'Diffx': -kx * z.imag, 'Diffy': -ky * z.imag, 'Diffxx': -kx ** 2 * z.real,
           'Diffyy': -ky ** 2 * z.real, 'Diffxy': -kx * ky * z.real}
    for name, e in exp.items():
        scale = np.max(np.abs(e)) + 1e-30
        assert np.max(np.abs(f[name] - e)) < 1e-6 * scale, name


def test_single_mode_choppy_and_current():
    Nx, Ny, dx, dy = 32, 32, 1.0, 1.0
    c = np.zeros((Ny, Nx), complex)
    c[3, 2] = 40.0 + 10.0j
    s = _surf(Nx, Ny, dx, dy, c, choppy=True, current=(0.5, -0.25))
    kx, ky = float(s.kx_s[2]), float(s.ky_s[3])
    k = np.hypot(kx, ky)
    t = 1.1
    f = osurf.synth(s, t, ('D', 'Diff'))
    n = np.arange(Nx)[None, :]
    m = np.arange(Ny)[:, None]
    z = complex(np.complex64(c[3, 2])) * np.exp(-1j * np.sqrt(G * k) * t) / (Nx * Ny) * \
        np.exp(2j * np.pi * (2 * n / Nx + 3 * m / Ny))
    np.testing.assert_allclose(f['Diffx'], -kx * z.imag / (1 - kx ** 2 / k * z.real),
                               atol=1e-6 * abs(kx) * abs(z).max())
    np.testing.assert_allclose(f['Diffy'], -ky * z.imag / (1 - ky ** 2 / k * z.real),
                               atol=1e-6 * abs(ky) * abs(z).max())
    # horizontal displacement carries the current drift current * t
    np.testing.assert_allclose(f['Dx'], -(kx / k) * z.imag + 0.5 * t, atol=1e-6)
    np.testing.assert_allclose(f['Dy'], -(ky / k) * z.imag - 0.25 * t, atol=1e-6)


def test_parseval_normalisation():
    rng = np.random.default_rng(3)
    Nx = Ny = 64
    c = (rng.standard_normal((Ny, Nx)) + 1j * rng.standard_normal((Ny, Nx))) * 10
    s = _surf(Nx, Ny, 1.0, 1.0, c)
    f = osurf.synth(s, 0.0, ('D',))
    cc = s.wave_coefs.astype(complex)
    # Dz = Re ifft2(c) = ifft2(H(c)): sum Dz^2 = sum |H(c)|^2 / (Nx Ny)
    ci = np.roll(np.flip(cc, (0, 1)), 1, (0, 1))          # c[-k]
    h = 0.5 * (cc + np.conj(ci))
    assert np.isclose(np.sum(f['Dz'].astype(float) ** 2), np.sum(np.abs(h) ** 2) / (Nx * Ny),
                      rtol=1e-6)
