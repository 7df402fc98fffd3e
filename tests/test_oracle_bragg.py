"""Pinning the oracle's Bragg RCS (romeiser97.py:251-393).

Checks: the flat-surface limit reduces to the first-order small-perturbation
transfer function 8 pi k0^4 cos^4(theta) |g_pq|^2 with Romeiser's g_hh, g_vv
and k_b = 2 k0 sin(theta); the full RCS equals T S(k_b)/k_b G(theta) dx dy with
the spectrum interpolated by scipy's own interp1d; slope-mirror symmetry; the
bad-facet mask (including the DP hh quirk); the spreading half plane.
"""
import numpy as np
from scipy.interpolate import interp1d

import oracle.bragg as ob
from oceansar_amd.spec import elfouhaily

K0 = 2 * np.pi * 5.405e9 / 299792458.0
EPS = 73 + 18j


def _table(U=10.):
    k = np.arange(1, 2 * K0, 0.1)
    return k, elfouhaily(k, U, 500e3)


def test_interp_matches_scipy_interp1d():
    k, S = _table()
    x = np.concatenate([np.linspace(0.5, 2 * K0 + 3, 5001), k[::50], [0.0, k[-1], k[0]]])
    ref = interp1d(k, S, bounds_error=False, fill_value=1)(x)
    np.testing.assert_allclose(ob.interp_fill1(x, k, S), ref, rtol=1e-14, atol=0)


def test_flat_surface_is_spm():
    theta = np.deg2rad(np.array([[25., 35., 45.]]))
    az = np.array([[0.0]])
    z = np.zeros((1, 3))
    k_b, th, T_hh, T_vv, bad = ob.facet_terms(theta, az, z, z, K0, 0.3, 0.125)
    np.testing.assert_allclose(k_b, 2 * K0 * np.sin(theta), rtol=1e-14)
    c, s2 = np.cos(theta), np.sin(theta) ** 2
    g_hh = EPS / (c + np.sqrt(EPS)) ** 2
    g_vv = EPS ** 2 * (1 + s2) / (EPS * c + np.sqrt(EPS)) ** 2
    np.testing.assert_allclose(T_hh, 8 * np.pi * K0 ** 4 * c ** 4 * np.abs(g_hh) ** 2, rtol=1e-12)
    np.testing.assert_allclose(T_vv, 8 * np.pi * K0 ** 4 * c ** 4 * np.abs(g_vv) ** 2, rtol=1e-12)
    np.testing.assert_allclose(th[0], -0.3, atol=1e-15)          # phi_b = az, minus wind
    np.testing.assert_allclose(th[1], np.pi - 0.3, atol=1e-15)
    assert not bad.any()


def test_flat_rcs_closed_form():
    k, S = _table(8.)
    inc = np.deg2rad(np.linspace(30., 40., 7))
    z = np.zeros((4, 7))
    az = np.array([0.0, 0.004, -0.01, 0.02])
    out = ob.rcs_romeiser97(inc, az, z, z, K0, 'vv', 0.2, 0.125, k, S, 0, 8., 500e3,
                            np.float32(0.5), np.float32(0.5))['vv']
    kb = 2 * K0 * np.sin(inc)[None, :]
    c = np.cos(inc)[None, :]
    s2 = np.sin(inc)[None, :] ** 2
    T = 8 * np.pi * K0 ** 4 * c ** 4 * np.abs(EPS ** 2 * (1 + s2) / (EPS * c + np.sqrt(EPS)) ** 2) ** 2
    S1 = np.interp(kb, k, S)
    for comp, th in ((0, az[:, None] - 0.2), (1, az[:, None] + np.pi - 0.2)):
        th = np.angle(np.exp(1j * th))
        G = ob.spread_elfouhaily(kb, th, 8., 500e3)
        np.testing.assert_allclose(out[comp], T * S1 / kb * G * 0.25, rtol=1e-12)
    assert np.all(out[1] == 0) and np.all(out[0] > 0)    # downwind component outside half plane


def test_slope_mirror_symmetry():
    rng = np.random.default_rng(5)
    k, S = _table()
    dxs = rng.normal(0, 0.1, (16, 16))
    dys = rng.normal(0, 0.1, (16, 16))
    inc = np.full(16, np.deg2rad(35.))
    az = np.zeros(16)
    a = ob.rcs_romeiser97(inc, az, dxs, dys, K0, 'vv', 0.0, 0.125, k, S, 0, 10., 500e3, 1., 1.)
    b = ob.rcs_romeiser97(inc, az, dxs, -dys, K0, 'vv', 0.0, 0.125, k, S, 0, 10., 500e3, 1., 1.)
    np.testing.assert_allclose(a['vv'], b['vv'], rtol=1e-12, atol=1e-300)


def test_bad_mask_and_dp_quirk():
    k, S = _table()
    inc = np.full(3, np.deg2rad(35.))
    az = np.zeros(1)
    # s_p close to the incidence angle: 4 sin^2(theta - s_p) < d^2
    dxs = np.array([[np.tan(np.deg2rad(35.)), 0.0, np.tan(np.deg2rad(34.))]])
    z = np.zeros((1, 3))
    with np.errstate(invalid='ignore', divide='ignore'):
        out = ob.rcs_romeiser97(inc, az, dxs, z, K0, 'DP', 0.0, 0.125, k, S, 0, 10., 500e3, 1., 1.)
    assert out['vv'][0, 0, 0] == 0 and out['vv'][1, 0, 0] == 0
    assert out['vv'][0, 0, 2] == 0 and out['vv'][1, 0, 2] == 0
    assert out['hh'][1, 0, 0] == 0 and out['hh'][1, 0, 2] == 0
    assert out['hh'][0, 0, 2] != 0          # DP: hh minus component is not masked
    assert out['vv'][1, 0, 1] != 0 or out['vv'][0, 0, 1] != 0


def test_cos2s_half_plane_and_normalisation():
    th = np.linspace(-np.pi, np.pi, 200001)
    G = ob.spread_cos2s(None, th, 0, 0)
    assert np.all(G[np.abs(th) > np.pi / 2] == 0)
    assert np.isclose(np.trapezoid(G, th), 0.5, rtol=1e-6)     # the reference's 1/2
    Ge = ob.spread_elfouhaily(np.full_like(th, 50.), th, 10., 500e3)
    assert np.isclose(np.trapezoid(Ge, th), 1.0, rtol=1e-4)
