"""CPU pins of the InSAR / ATI oracle (oracle/insar.py) and of the host side of
insar_process / ati_process.

  - smooth1d and smooth (misc.py:108-197): brute-force window sums, the edge
    count normalisation, fftconvolve 'same' centring for every window type,
    the identity below length 3;
  - the azimuth shift (insar_processor.py:165-174): an integer shift is a roll,
    a tone is an eigenvector whose eigenvalue is the closed-form phasor at the
    reference's rolled frequency index, including the half-bin offset the roll
    gives an odd number of lines;
  - the device algorithm of the shift, run with numpy FFTs: the wrapped-kernel
    linear convolution (and its two-half split) reproduces the oracle;
  - sr_to_geo (geometry.py:101-133): golden vectors from the reference's own
    source text, bit for bit, and its incidence against inc_to_sr;
  - coherence: identical images give 1 and phase 0, a constant phase offset is
    recovered, a planted pair-order error is detected;
  - uwphase: the centre is the left edge of the fullest bin (a point just
    beyond centre + pi wraps down);
  - the L1b file layout round trip.
"""
import os

import numpy as np
import pytest

import oracle.insar as oi

HERE = os.path.dirname(os.path.abspath(__file__))


def brute_flat(x, L, axis):
    x = np.moveaxis(x, axis, 0)
    n = x.shape[0]
    out = np.zeros_like(x)
    for i in range(n):
        acc = np.zeros_like(x[0])
        cnt = 0
        for d in range(-(L // 2), L - L // 2):
            if 0 <= i + d < n:
                acc = acc + x[i + d]
                cnt += 1
        out[i] = acc / cnt
    return np.moveaxis(out, 0, axis)


def brute_taps(x, w, axis):
    """out[n] = sum_t w[t] x[n + (L - 1) // 2 - t], zero outside."""
    x = np.moveaxis(x, axis, 0)
    n, L = x.shape[0], len(w)
    out = np.zeros(x.shape, dtype=np.result_type(x, w))
    for i in range(n):
        for t in range(L):
            j = i + (L - 1) // 2 - t
            if 0 <= j < n:
                out[i] = out[i] + w[t] * x[j]
    return np.moveaxis(out, 0, axis)


@pytest.mark.parametrize('L', [1, 2, 3, 4, 5, 8, 9])
@pytest.mark.parametrize('axis', [0, 1])
def test_smooth1d_is_the_edge_normalised_moving_average(L, axis):
    rng = np.random.default_rng(L + 10 * axis)
    x = rng.standard_normal((13, 11)) + 1j * rng.standard_normal((13, 11))
    got = oi.smooth(x, L, axis=axis)
    np.testing.assert_allclose(got, brute_flat(x, L, axis), rtol=1e-14, atol=1e-15)
    # the sums run in offset order from zero: bit-identical to that order
    np.testing.assert_array_equal(oi.smooth1d(x.real, L, axis), brute_flat(x.real, L, axis))


@pytest.mark.parametrize('window', ['hanning', 'hamming', 'bartlett', 'blackman'])
@pytest.mark.parametrize('L', [3, 4, 5, 8])
def test_smooth_taps_is_centred_zero_padded_correlation(window, L):
    rng = np.random.default_rng(L)
    x = rng.standard_normal((17, 9)) + 1j * rng.standard_normal((17, 9))
    w = getattr(np, window)(L)
    w = w / w.sum()
    for axis in (0, 1):
        np.testing.assert_allclose(oi.smooth(x, L, window=window, axis=axis),
                                   brute_taps(x, w, axis), rtol=1e-12, atol=1e-13)


def test_smooth_short_windows_and_constants():
    x = np.arange(12.).reshape(3, 4)
    assert oi.smooth(x, 2, window='hanning', axis=1) is x
    c = np.full((7, 5), 3.25)
    np.testing.assert_array_equal(oi.smooth(c, 4, axis=0), c)   # count-normalised: exact
    with pytest.raises(ValueError):
        oi.smooth(x, 3, window='kaiser', axis=0)


def _rolled_index(k, N):
    """The reference's frequency index after np.roll(arange(N) - N/2, int(-N/2)),
    in closed form: k for k < ceil(N/2) else k - N, minus 1/2 when N is odd."""
    half = 0.5 if N % 2 else 0.
    return (k - half) if k < (N + 1) // 2 else (k - N - half)


@pytest.mark.parametrize('N', [8, 9, 64, 77])
def test_shift_phasor_closed_form(N):
    d = 0.37
    S = oi.shift_phasor(N, d)
    want = np.exp(-2j * np.pi * d * np.array([_rolled_index(k, N) for k in range(N)]) / N)
    np.testing.assert_allclose(S, want, rtol=0, atol=1e-14)


@pytest.mark.parametrize('N', [16, 40])
def test_integer_shift_is_a_roll(N):
    rng = np.random.default_rng(N)
    x = (rng.standard_normal((1, 1, N, 3)) + 1j * rng.standard_normal((1, 1, N, 3))
         ).astype(np.complex64)
    v, prf = 7000., 1750.                         # v / prf = 4 m per line
    got = oi.coregister(x.copy(), [-8.], v, prf)  # shift_dist = +2 lines
    np.testing.assert_allclose(got, np.roll(x, 2, axis=2), rtol=0, atol=1e-6)


@pytest.mark.parametrize('N', [32, 33])
def test_tone_is_an_eigenvector_of_the_shift(N):
    k0 = 5
    n = np.arange(N)
    x = np.exp(2j * np.pi * k0 * n / N).astype(np.complex64)[None, None, :, None]
    b, v, prf = 1.3, 7000., 1750.
    got = oi.coregister(x.copy(), [b], v, prf)[0, 0, :, 0]
    dist = -b / (v / prf)
    lam = np.exp(-2j * np.pi * dist * _rolled_index(k0, N) / N)
    np.testing.assert_allclose(got, (x[0, 0, :, 0] * lam).astype(np.complex64), atol=2e-7)


def _device_algorithm(x, shift_dist):
    """The device's convolution, with numpy FFTs, for a column x (complex128)."""
    from oceansar_amd.insar_processor import shift_kernel
    N = len(x)
    M, G = shift_kernel(N, shift_dist)
    if M <= 8192:
        return np.fft.ifft(np.fft.fft(x, M) * G)[:N] * M
    H = M // 2
    xp = np.zeros(H, complex)
    xp[:N] = x
    n = np.arange(H)
    E = np.fft.fft(xp) * G[0::2]
    O = np.fft.fft(xp * np.exp(-2j * np.pi * n / M)) * G[1::2]
    U, V = np.fft.ifft(E) * H, np.fft.ifft(O) * H
    return (U + np.exp(2j * np.pi * n / M) * V)[:N]


@pytest.mark.parametrize('N', [1, 2, 7, 64, 4095, 4097, 7555])
def test_wrapped_kernel_convolution_reproduces_the_shift(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    d = -0.731
    want = np.fft.ifft(np.fft.fft(x) * oi.shift_phasor(N, d))
    got = _device_algorithm(x, d)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-13


def test_sr_to_geo_golden_and_closed_form():
    from oceansar_amd.geometry import inc_to_sr, sr_to_geo
    g = np.load(os.path.join(HERE, 'golden', 'golden_insar.npz'))
    for n in ('115', '267', '624'):
        out = sr_to_geo(g['srgeo_sr_' + n], 514e3)
        for k, o in zip(('gr', 'inc', 'look', 'b'), out):
            np.testing.assert_array_equal(o, g['srgeo_%s_%s' % (k, n)])
    inc = np.deg2rad(np.array([20., 35., 45.]))
    _, got, _, _ = sr_to_geo(inc_to_sr(inc, 514e3), 514e3)
    np.testing.assert_allclose(got, inc, rtol=0, atol=1e-8)


def _pair_images(C, shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((C,) + shape) + 1j * rng.standard_normal((C,) + shape)


def test_coherence_of_identical_and_phase_shifted_images():
    x = _pair_images(1, (40, 30), 1)[0]
    phi = 0.8
    data = np.stack([x, x, x * np.exp(1j * phi)])[:, None]      # 3 channels, 1 pol
    i_all, cohs, lut = oi.l1b(data, 5, 5, 'hanning', 5, 35, 0, 30)
    np.testing.assert_allclose(np.abs(cohs), 1., atol=1e-13)
    np.testing.assert_allclose(np.angle(cohs[lut[0, 1]]), 0., atol=1e-13)
    np.testing.assert_allclose(np.angle(cohs[lut[0, 2]]), phi, atol=1e-13)
    np.testing.assert_allclose(np.angle(cohs[lut[1, 2]]), phi, atol=1e-13)
    assert i_all.shape == (3, 1, 30, 30)
    # planted mistake: the conjugate on the other image flips the sign
    t = oi.smooth(oi.smooth(data[0, 0] * np.conj(data[2, 0]), 5, axis=1, window='hanning'),
                  5, axis=0, window='hanning')[5:35]
    assert np.allclose(np.angle(t), -phi)


def test_coherence_lut_order_and_bounds():
    data = _pair_images(4, (20, 16), 2)[:, None]
    i_all, cohs, lut = oi.l1b(data, 3, 3, 'flat', 0, 20, 0, 16)
    assert [lut[a, b] for a in range(4) for b in range(a + 1, 4)] == list(range(6))
    assert (np.abs(cohs) <= 1 + 1e-12).all()
    unit = np.exp(1j * np.angle(data))
    i1, _, _ = oi.l1b(unit, 4, 2, 'flat', 0, 20, 0, 16)
    np.testing.assert_allclose(i1, 1., rtol=1e-15)


def test_uwphase_centre_is_left_bin_edge():
    nb = 72
    w = 2 * np.pi / nb
    e10 = -np.pi + 10 * w
    ph = np.concatenate([e10 + w * np.linspace(0.1, 0.9, 3590), [e10 + np.pi + w / 4]])
    ph = np.angle(np.exp(1j * ph))          # wrapped to (-pi, pi]
    got = oi.uwphase(np.exp(1j * ph))
    # every value lies in (e10 - pi, e10 + pi]: the extra point e10 + pi + w/4 wraps down
    assert got[-1] == pytest.approx(e10 + np.pi + w / 4 - 2 * np.pi, abs=1e-12)
    np.testing.assert_allclose(got[:-1], ph[:-1], atol=1e-12)
    assert (got > e10 - np.pi - 1e-12).all() and (got <= e10 + np.pi + 1e-12).all()


def test_window_spec_and_slice_range():
    from oceansar_amd.insar_processor import slice_range, window_spec
    assert window_spec('flat', 5.9) == (5, 1, None)
    L, f, t = window_spec('hanning', 2)
    assert (L, f) == (1, 0) and t.tolist() == [1.]
    L, f, t = window_spec('hamming', 6)
    np.testing.assert_allclose(t, np.hamming(6) / np.hamming(6).sum())
    assert slice_range(3, 10, 12, 'az') == (3, 10)
    with pytest.raises(ValueError):
        slice_range(-2, 10, 12, 'az')        # numpy would take [10, 10): empty, not 12 lines
    with pytest.raises(ValueError):
        slice_range(0, 14, 12, 'rg')


def test_l1b_file_round_trip(tmp_path):
    from oceansar_amd import raw_io
    rng = np.random.default_rng(3)
    i_all = rng.random((2, 1, 5, 4))
    coh = rng.random((1, 5, 4))
    pha = rng.random((1, 5, 4))
    p = str(tmp_path / 'l1b.nc')
    with raw_io.L1bFile(p, 'w', i_all.shape) as f:
        f.set('ml_intensity', i_all)
        f.set('ml_coherence', coh)
        f.set('ml_phase', pha)
        f.set('coh_lut', np.array([[-1, 0], [-1, -1]]).reshape(2, 1, 2, 1))
        f.set('b_ati', [0., 10.])
        f.set('rg_ml', 5)
        f.set('inc_angle', 0.61)
    with raw_io.L1bFile(p, 'r') as f:
        np.testing.assert_array_equal(f.get('ml_intensity'), i_all)
        np.testing.assert_array_equal(f.get('ml_phase'), pha)
        assert f.get('coh_lut')[0, 0, 1, 0] == 0 and f.get('rg_ml') == 5
        assert f._f.variables['ml_coherence'].dimensions == ('covel_dim', 'az_dim', 'rg_dim')
        assert f._f.variables['ml_intensity'].dimensions == ('ch_dim', 'pol_dim', 'az_dim',
                                                             'rg_dim')
