"""Pinning the oracle's SAR focusing (sar_processor.py:131-229) with point targets.

A range-compressed point target with the exact hyperbolic range history
R(t) = sqrt(R0^2 + v^2 (t - t0)^2), phase exp(-4 pi i (R(t) - sr0) / lambda),
the sampled range response sinc((n - (R(t) - sr0) / dr) / over_fs) and the
two-way antenna pattern of the processor, must focus to a compact peak at its
closest-approach sample t0 prf and range sample (R0 - sr0) / dr, with the
phase -4 pi (R0 - sr0) / lambda - pi / 4. A wrong sign of the azimuth matched filter
or of the range-migration phase, or frequencies in the wrong order, spreads the
peak or moves it.
"""
import numpy as np
import pytest

import oracle.focus as of

C = 299792458.0
F0, PRF, V, SR0 = 5.405e9, 1800., 7600., 616400.5
L0 = C / F0
FS, OVER = 110e6, 1.1
DR = C / (2 * FS)
ANT = 6.0


def _target(n_az, n_rg, a0, x0, amp=1.0):
    """Raw [n_az][n_rg] of one point target at azimuth sample a0, range sample x0."""
    t = (np.arange(n_az) - a0) / PRF
    R0 = SR0 + x0 * DR
    R = np.sqrt(R0 ** 2 + (V * t) ** 2)
    sin_az = -V * t / R
    beam = of.sinc_bp(sin_az, ANT, F0, field=True) ** 2
    n = np.arange(n_rg)
    x = (R - SR0) / DR
    rng = np.sinc((n[None, :] - x[:, None]) / OVER)
    return amp * beam[:, None] * rng * np.exp(-4j * np.pi * (R - SR0) / L0)[:, None]


def _focus(raw, weighting=1.0, doppler_bw=1200.):
    return of.focus_channel(raw[None], PRF, FS, F0, V, SR0, doppler_bw, weighting, ANT, ANT)[0]


def _lo(n_az, doppler_bw=1200.):
    return of.valid_rows(n_az, doppler_bw, V, F0, SR0, PRF)[0]


@pytest.mark.parametrize('weighting', [1.0, 0.54])
def test_point_target_focuses_at_its_position_and_phase(weighting):
    n_az, n_rg = 2048, 128
    a0, x0 = 1000, 40.3
    slc = _focus(_target(n_az, n_rg, a0, x0), weighting)
    lo = _lo(n_az)
    mag = np.abs(slc)
    ia, ir = np.unravel_index(np.argmax(mag), mag.shape)
    assert ia + lo == a0 and ir == round(x0)
    # sub-sample range position: the focused range response is the target's
    # sinc((n - x0) / over_fs), so the two samples around x0 have its ratio
    ratio = mag[ia, ir + 1] / mag[ia, ir]
    want_ratio = abs(np.sinc((ir + 1 - x0) / OVER) / np.sinc((ir - x0) / OVER))
    assert abs(ratio - want_ratio) < 0.02, (ratio, want_ratio)
    # compact: most of the energy within +-2 azimuth and range samples of the peak
    e = np.sum(mag ** 2)
    e_peak = np.sum(mag[ia - 2:ia + 3, ir - 2:ir + 3] ** 2)
    assert e_peak / e > 0.8, e_peak / e
    # phase of the peak: the two-way path to closest approach, and -pi/4: the
    # azimuth chirp's spectrum has the stationary-phase factor exp(-i pi/4),
    # which the reference's filter exp(i ph_ac) leaves in place
    want = -4 * np.pi * (x0 * DR) / L0 - np.pi / 4
    got = np.angle(slc[ia, round(x0)] * np.exp(-1j * want))
    assert abs(got) < 0.05, got


def test_two_targets_keep_their_separation():
    n_az, n_rg = 2048, 128
    raw = _target(n_az, n_rg, 700, 20.0) + _target(n_az, n_rg, 1320, 71.0, amp=0.5)
    slc = _focus(raw)
    lo = _lo(n_az)
    mag = np.abs(slc)
    a1, r1 = np.unravel_index(np.argmax(mag), mag.shape)
    assert (a1 + lo, r1) == (700, 20)
    mag2 = mag.copy()
    mag2[a1 - 8:a1 + 9, :] = 0
    a2, r2 = np.unravel_index(np.argmax(mag2), mag2.shape)
    assert (a2 + lo, r2) == (1320, 71)
    assert abs(mag[a2, r2] / mag[a1, r1] - 0.5) < 0.02        # linearity


def test_focusing_is_linear_and_shift_covariant():
    rng = np.random.default_rng(3)
    raw = rng.standard_normal((2, 1024, 64)) + 1j * rng.standard_normal((2, 1024, 64))
    a = of.focus_channel(raw, PRF, FS, F0, V, SR0, 1200., 1., ANT, ANT)
    b = of.focus_channel(2.5j * raw, PRF, FS, F0, V, SR0, 1200., 1., ANT, ANT)
    np.testing.assert_allclose(b, 2.5j * a, rtol=1e-12, atol=1e-12 * np.abs(a).max())
    # circular shift along azimuth by k samples shifts the (uncropped) image by k
    lo = _lo(1024)
    k = 16
    c = of.focus_channel(np.roll(raw, k, axis=1), PRF, FS, F0, V, SR0, 1200., 1., ANT, ANT)
    np.testing.assert_allclose(c[:, k:], a[:, :-k], rtol=0, atol=1e-9 * np.abs(a).max())
    assert lo > 0


def test_tables_match_the_reference_formulas():
    """fftfreq ordering, the Doppler-domain range migration and the flat
    weighting's normalisation to unit mean power over the kept band."""
    fr, fa, rcmc, h = of.tables(1024, 128, PRF, FS, F0, V, SR0, 1200., 1.0, ANT, ANT)
    assert fa[1] == PRF / 1024 and fa[-1] == -PRF / 1024 and fr[1] == FS / 128
    np.testing.assert_allclose(rcmc, SR0 / np.sqrt(1 - (fa * L0 / 2 / V) ** 2) - SR0, rtol=1e-15)
    band = np.abs(h) > 0
    assert band.sum() == int(1200. / (PRF / 1024))
    assert np.all(np.abs(fa[band]) <= 600. + PRF / 1024)
