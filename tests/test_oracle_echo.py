"""Pinning the oracle's range integrator (range_profile.py:15-49,83-116).

The reference has no tests; these checks fix the integrator by brute force
(an independent sinc evaluation with the same 1/20-bin quantisation), a point
target, edge dropping, linearity, a one-bin shift and round-half-to-even.
"""
import numpy as np

import oracle.echo as oecho

FS = 1.1
TAB = oecho.calc_sinc_vec(8, 20, Fs=FS)


def brute(srg, scene, binsize, min_sr, nbins):
    """out[b] = sum_i scene_i * sinc((b + 1 - xq_i)/Fs) over the 8 bins b in [d-4, d+3],
    xq = d + 1 - round(20 (1 - frac))/20, evaluated with np.sinc directly."""
    out = np.zeros(nbins, dtype=np.complex128)
    for s, v in zip(srg, scene):
        x = (s - min_sr) / binsize
        d = int(np.floor(x))
        sb = int(np.round(20 * (1 - (x - d))))
        for b in range(d - 4, d + 4):
            if 0 <= b < nbins:
                w = np.float32(np.sinc(((b - d) + sb / 20.) / FS))
                out[b] += float(w) * v
    return out


def test_integrator_matches_brute_force():
    rng = np.random.default_rng(1)
    n, nbins, bs, mn = 3000, 120, 1.37, -2.73
    srg = rng.uniform(mn - 8 * bs, mn + (nbins + 8) * bs, n)
    sc = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    out = np.zeros(nbins, dtype=np.complex128)
    oecho.chan_profile(srg, sc, bs, mn, TAB, 8, 20, out)
    ref = brute(srg, sc, bs, mn, nbins)
    assert np.max(np.abs(out - ref)) < 1e-5 * np.max(np.abs(ref))


def test_point_target_taps():
    out = np.zeros(100, dtype=np.complex128)
    oecho.chan_profile(np.array([50.5]), np.array([1 + 0j]), 1.0, 0.0, TAB, 8, 20, out)
    nz = np.nonzero(out)[0]
    assert list(nz) == list(range(46, 54))
    # d = 50, frac = 0.5 -> sb = 10 -> weights sinc((j - 4 + 0.5)/Fs)
    expect = np.sinc((np.arange(8) - 3.5) / FS)
    np.testing.assert_allclose(out[46:54].real, expect, rtol=2e-7)
    # the peak sits one bin late (compensated by sr_near in the driver)
    assert np.argmax(np.abs(out)) in (49, 50)


def test_edge_taps_are_dropped():
    for x, keep in ((1.2, range(0, 5)), (97.6, range(93, 100))):
        out = np.zeros(100, dtype=np.complex128)
        oecho.chan_profile(np.array([x]), np.array([1 + 0j]), 1.0, 0.0, TAB, 8, 20, out)
        assert set(np.nonzero(out)[0]) == set(keep)


def test_linearity_and_shift():
    rng = np.random.default_rng(2)
    srg = rng.uniform(10, 80, 500)
    a = rng.standard_normal(500) + 1j * rng.standard_normal(500)
    b = rng.standard_normal(500) + 1j * rng.standard_normal(500)
    oa, ob, oab = (np.zeros(120, complex) for _ in range(3))
    oecho.chan_profile(srg, a, 1.0, 0.0, TAB, 8, 20, oa)
    oecho.chan_profile(srg, b, 1.0, 0.0, TAB, 8, 20, ob)
    oecho.chan_profile(srg, 2 * a - 3j * b, 1.0, 0.0, TAB, 8, 20, oab)
    np.testing.assert_allclose(oab, 2 * oa - 3j * ob, atol=1e-12)
    sh = np.zeros(120, complex)
    # a shift by exactly one bin (binary-exact offsets) moves the profile by one bin
    srg_q = np.round(srg * 64) / 64
    o1, o2 = np.zeros(120, complex), np.zeros(120, complex)
    oecho.chan_profile(srg_q, a, 1.0, 0.0, TAB, 8, 20, o1)
    oecho.chan_profile(srg_q + 1.0, a, 1.0, 0.0, TAB, 8, 20, o2)
    np.testing.assert_allclose(o2[1:], o1[:-1], atol=1e-12)
    del sh


def test_subbin_round_half_to_even():
    # binary-exact fractions: frac = 0.875 -> 20 * 0.125 = 2.5 -> 2 (even);
    # frac = 0.625 -> 7.5 -> 8 (even)
    d, sb = oecho.bin_index(np.array([3.875]), 1.0, 0.0)
    assert d[0] == 3 and sb[0] == 2
    d, sb = oecho.bin_index(np.array([3.625]), 1.0, 0.0)
    assert sb[0] == 8


def test_lerp_is_balancer_float32():
    """Between snapshots the fields follow the reference balancer's float32
    interpolation (balancer.py:274-282): float32 weights, float32 products and
    sum. A case where FP64 evaluation rounds differently pins the precision."""
    import types
    import oracle.sar as osar
    a = np.array([[1 + 2 ** -23, 3.0]], dtype=np.float32)
    b = np.array([[1.0, np.float32(1 / 3)]], dtype=np.float32)
    snaps = {0: {'Dz': a}, 1: {'Dz': b}}
    sc = types.SimpleNamespace(K=3)
    cache = types.SimpleNamespace(get=snaps.get)
    got = osar.lerp_fields(cache, sc, 1, ('Dz',))['Dz']
    wh = np.float32(1 / 3)
    wl = np.float32(1 - wh)
    want = (np.float32(wl * a) + np.float32(wh * b)).astype(np.float64)
    fp64 = (1 - 1 / 3) * a.astype(np.float64) + (1 / 3) * b.astype(np.float64)
    assert got.dtype == np.float64
    assert np.array_equal(got, want)
    assert not np.array_equal(got, fp64)
    # at a snapshot pulse the snapshot itself
    assert np.array_equal(osar.lerp_fields(cache, sc, 3, ('Dz',))['Dz'], b.astype(np.float64))
