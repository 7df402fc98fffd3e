"""The counter-based draws and the coefficient build of the device surface
initialisation (oracle/surface_init.py).

Pins: Philox4x32-10 against the Random123 known-answer vectors (zero,
all-ones and pi-digit counters and keys); the draws are standard complex
Gaussians (moments, independence of the two streams and of neighbouring
points); the deterministic part equals the host OceanSurface.init restatement
(wave_dirspec to the bit); with one seed the coefficients follow the
reference's |c|^2 = 2 (Nx Ny)^2 dirspec dkx dky law on average.
"""
import numpy as np
import pytest

import oracle.surface_init as oi
from oceansar_amd.surface_init import init_surface


@pytest.mark.parametrize('ctr,key,want', [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, want):
    got = oi.philox4x32_10(*[np.array([c], np.uint64) for c in ctr], *key)
    assert tuple(int(g[0]) for g in got) == want


def test_draws_are_independent_standard_complex_gaussians():
    a, b = oi.draws(12345, 256, 256)
    for z in (a, b):
        z = z.astype(np.complex128).ravel()
        assert abs(np.mean(z)) < 0.01
        assert abs(np.mean(np.abs(z) ** 2) - 1) < 0.01
        assert abs(np.mean(z.real ** 2) - 0.5) < 0.01 and abs(np.mean(z.real * z.imag)) < 0.01
        # Rayleigh amplitude: E|z|^4 = 2 E|z|^2^2 for a circular Gaussian
        assert abs(np.mean(np.abs(z) ** 4) - 2) < 0.05
    assert abs(np.mean(a.ravel() * np.conj(b.ravel()))) < 0.01
    assert abs(np.mean(a.ravel()[1:] * np.conj(a.ravel()[:-1]))) < 0.01
    c, _ = oi.draws(12346, 256, 256)
    assert abs(np.mean(a.ravel() * np.conj(c.ravel()))) < 0.01
    np.testing.assert_array_equal(oi.draws(12345, 256, 256)[0], a)


@pytest.mark.parametrize('spread', ['elfouhaily', 'cos2s'])
def test_dirspec_matches_host_init(spread):
    rng = np.random.default_rng(0)
    s = init_surface(256., 128., 1., 1., 'elfouhaily', spread, np.deg2rad(30.), 500e3, 9.,
                     0.3, 1.2, rng_wave=rng)
    coefs, dirspec = oi.coefficients(s, 7, spread_model=spread)
    np.testing.assert_array_equal(dirspec, s.wave_dirspec)
    assert coefs.dtype == np.complex64 and dirspec.dtype == np.float32
    # E |c|^2 = 2 (Nx Ny)^2 dirspec dkx dky (ocean.py:236 with E|cg|^2 = 1)
    want = 2. * (s.Nx * s.Ny) ** 2 * dirspec.astype(float) * float(s.kx_res) * float(s.ky_res)
    assert abs(np.sum(np.abs(coefs.astype(complex)) ** 2) / np.sum(want) - 1) < 0.05
