"""The oracle's Bragg facet model, Elfouhaily spreading and facet range against
golden vectors written by the reference's own source text, plus the closed
forms of single-axis tilts.

tests/golden/golden_bragg.npz was produced by tests/golden/make_golden_bragg.py,
which executed RCSRomeiser97._rcs_numba (oceansar/nrcs/romeiser97.py:251-322),
spread.elfouhaily (oceansar/spread/elfouhaily.py:23-56) and the sr_surface /
wave_dinc statements of sar_raw (oceansar/radarsim/sar_raw_nompi.py:366-369) on
tilted facets (slopes of sigma 0.1 to 0.25 and a steep tail, incidence 20-50
deg, azimuth angles to 0.5 rad, three wind directions).
"""
import os

import numpy as np
import pytest

import oracle.bragg as ob
import oracle.sar as osar
from oceansar_amd import spread as host_spread

G = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'golden_bragg.npz'))
K0 = float(G['k0'][0])
EPS = 73 + 18j


@pytest.mark.parametrize('i', [0, 1, 2])
def test_facet_terms_match_reference(i):
    g = lambda k: G['rcs%d_%s' % (i, k)]
    with np.errstate(all='ignore'):
        k_b, th, T_hh, T_vv, bad = ob.facet_terms(g('inc'), g('az'), g('diffx'), g('diffy'), K0,
                                                  float(g('wind_dir')[0]), 0.125)
    np.testing.assert_allclose(k_b, g('k_b'), rtol=1e-13, atol=0)
    np.testing.assert_allclose(th, g('theta'), rtol=0, atol=1e-13)
    np.testing.assert_allclose(T_hh, g('T_hh'), rtol=1e-12, atol=0)
    np.testing.assert_allclose(T_vv, g('T_vv'), rtol=1e-12, atol=0)
    np.testing.assert_array_equal(bad, g('bad'))
    if i == 2:
        assert g('bad').any() and not g('bad').all()   # the steep set exercises the mask


def test_golden_sets_are_tilted():
    """The vectors exercise both tilt axes away from zero (what the flat tests miss)."""
    for i in range(3):
        az = G['rcs%d_az' % i]
        u = np.cos(az) * G['rcs%d_diffx' % i] + np.sin(az) * G['rcs%d_diffy' % i]
        v = -np.sin(az) * G['rcs%d_diffx' % i] + np.cos(az) * G['rcs%d_diffy' % i]
        assert np.std(np.arctan(u)) > 0.08 and np.std(np.arctan(v)) > 0.08


@pytest.mark.parametrize('U', [3., 7., 10., 12., 15.])
@pytest.mark.parametrize('fetch', [500e3, 50e3])
def test_elfouhaily_spreading_matches_reference(U, fetch):
    k, th = G['spread_k'], G['spread_th']
    ref = G['spread_U%g_F%g' % (U, fetch)]
    np.testing.assert_allclose(ob.spread_elfouhaily(k, th, U, fetch), ref, rtol=1e-13, atol=0)
    np.testing.assert_allclose(host_spread.elfouhaily(k, th, U, fetch), ref, rtol=1e-13, atol=0)
    # Delta(k) is visible: the angular shape is not the plain (1 + cos 2 theta)/pi
    up = np.abs(th) < np.pi / 2
    assert np.ptp(ref[up] * np.pi - 1 - np.cos(2 * th[up])) > 1e-3


def test_facet_range_matches_reference():
    sr_s, dinc = osar.facet_range(G['geo_sr'], np.cos(G['geo_inc']), np.sin(G['geo_inc']),
                                  G['geo_az'], G['geo_az'] / G['geo_sr0'][0], G['geo_Dx'],
                                  G['geo_Dy'], G['geo_Dz'], G['geo_sr0'][0])
    np.testing.assert_array_equal(sr_s, G['geo_sr_surface'])
    np.testing.assert_array_equal(dinc, G['geo_wave_dinc'])


def _spm(theta):
    c, s2 = np.cos(theta), np.sin(theta) ** 2
    g_hh = EPS / (c + np.sqrt(EPS)) ** 2
    g_vv = EPS ** 2 * (1 + s2) / (EPS * c + np.sqrt(EPS)) ** 2
    F = 8 * np.pi * K0 ** 4 * c ** 4
    return F * np.abs(g_hh) ** 2, F * np.abs(g_vv) ** 2


def test_pure_range_tilt_is_spm_at_local_incidence():
    """s_t = 0: the facet only tilts toward or away from the radar. The model is
    first-order SPM at the local incidence theta - s_p, weighted by the facet's
    projected area cos(theta - s_p)/(cos theta cos s_p), with k_b = 2 k0 sin(theta - s_p)
    pointing along the look direction."""
    inc = np.deg2rad(np.array([[30., 35., 40.]]))
    s_p = np.deg2rad(np.array([[-12.], [-4.], [3.], [9.]]))
    diffx = np.tan(s_p) + 0 * inc
    z = np.zeros_like(diffx)
    k_b, th, T_hh, T_vv, bad = ob.facet_terms(inc, np.zeros((4, 1)), diffx, z, K0, 0.0, 0.125)
    loc = inc - s_p
    w = np.cos(loc) / (np.cos(inc) * np.cos(s_p))
    Thh, Tvv = _spm(loc)
    np.testing.assert_allclose(k_b, 2 * K0 * np.sin(loc), rtol=1e-13)
    np.testing.assert_allclose(T_hh, w * Thh, rtol=1e-12)
    np.testing.assert_allclose(T_vv, w * Tvv, rtol=1e-12)
    np.testing.assert_allclose(th[0], 0 * loc, atol=1e-15)
    assert not bad.any()


def test_pure_azimuth_tilt_local_frame():
    """s_p = 0, az = 0: the facet tilts about the look direction by s_t. The
    local incidence is cos theta_i = cos theta cos s_t, the Bragg wave number is
    2 k0 times the projection of the incident unit vector on the facet plane,
    |proj| = sin theta_i, and its direction in the facet's tangent frame
    (e_x, e_y' = (0, cos s_t, sin s_t)) is atan2(proj . e_y', proj . e_x)."""
    inc = np.deg2rad(np.array([[30., 35., 40.]]))
    s_t = np.deg2rad(np.array([[-15.], [-5.], [6.], [14.]]))
    diffy = np.tan(s_t) + 0 * inc
    z = np.zeros_like(diffy)
    k_b, th, T_hh, T_vv, bad = ob.facet_terms(inc, np.zeros((4, 1)), z, diffy, K0, 0.0, 0.125)
    cti = np.cos(inc) * np.cos(s_t)
    np.testing.assert_allclose(k_b, 2 * K0 * np.sqrt(1 - cti ** 2), rtol=1e-13)
    # incident direction (sin theta, 0, -cos theta); facet normal (0, -sin s_t, cos s_t)
    n_i = np.stack([np.sin(inc) + 0 * s_t, 0 * inc * s_t, -np.cos(inc) + 0 * s_t])
    n = np.stack([0 * s_t + 0 * inc, -np.sin(s_t) + 0 * inc, np.cos(s_t) + 0 * inc])
    proj = n_i - np.sum(n_i * n, axis=0) * n
    np.testing.assert_allclose(np.sqrt(np.sum(proj ** 2, axis=0)), np.sqrt(1 - cti ** 2),
                               rtol=1e-13)
    e_y = np.stack([0 * s_t + 0 * inc, np.cos(s_t) + 0 * inc, np.sin(s_t) + 0 * inc])
    phi = np.arctan2(np.sum(proj * e_y, axis=0), proj[0])
    np.testing.assert_allclose(th[0], phi, atol=1e-13)
    # a = sin^2 s_t / sin^2 theta_i mixes hh into vv (and back): not plain SPM at theta_i
    Thh_i, Tvv_i = _spm(np.arccos(cti))
    assert np.all(np.abs(T_vv / Tvv_i - 1) > 1e-4)


def test_kb_beyond_2k0_gives_reference_non_finite():
    """The reference forces k_b > 2 k0 to 0 (romeiser97.py:307), then divides the
    fill value 1 of the table by it: the RCS is inf where G > 0 and NaN where
    G = 0, and the bad-facet mask still zeroes it."""
    k = np.arange(1, 2 * K0, 0.1)
    from oceansar_amd.spec import elfouhaily
    S = elfouhaily(k, 10., 500e3)
    k_b = np.array([[0.0, 0.0, 150.0, 0.0]])
    th = np.array([[[0.3, 2.0, 0.3, 0.3]], [[0.3 + np.pi - 2 * np.pi, 2.0 - np.pi, -2.8, 0.3]]])
    T = np.full((1, 4), 1e3)
    bad = np.array([[False, False, False, True]])
    for sid in (0, 1):
        out = ob.rcs_from_terms(k_b, th, T, T, bad, 'DP', k, S, sid, 10., 500e3, 1., 1.)
        vv, hh = out['vv'], out['hh']
        assert np.isinf(vv[0, 0, 0]) and vv[0, 0, 0] > 0       # G > 0 over k_b = 0
        assert np.isnan(vv[0, 0, 1])                            # |theta| > pi/2: G = 0
        assert np.isfinite(vv[:, 0, 2]).all()
        assert vv[0, 0, 3] == 0 and vv[1, 0, 3] == 0             # masked
        assert np.isinf(hh[0, 0, 3])                             # DP: hh minus not masked
