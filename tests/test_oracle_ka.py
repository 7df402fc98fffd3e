"""Pins of the Kirchhoff stationary-phase oracle (oracle/ka.py, restating
oceansar/nrcs/ka.py:131-270) against closed forms it must reduce to.

Geometrical optics: a doubly curved surface seen at its specular point
backscatters sigma = pi R1 R2 |R(0)|^2, with R1 R2 = 1/K_G the product of the
principal radii (K_G = J / (1 + |grad z|^2)^2 the Gaussian curvature, J the
Hessian determinant) and R(0) the normal-incidence Fresnel coefficient - a
sphere of radius a gives pi a^2 |R(0)|^2. The driver scales the field as
Esn = (4 pi)^1.5 E (sar_raw_nompi.py:405-426), so |Esn|^2 must equal sigma.
This checks the constant K, the 2 pi / q_z and n_norm / sqrt|J| factors and
the polarisation vector magnitude together. The stationary-point offset enters
as exp(i q . delta): moving the stationary point by delta changes the phase by
q . delta.
"""
import numpy as np
import pytest

from oracle.ka import EPS_SW, field_spa

K0 = 2 * np.pi * 5.405e9 / 299792458.
R0 = (1 - np.sqrt(EPS_SW)) / (1 + np.sqrt(EPS_SW))


def _facet(th, ph, Dxx, Dyy, Dxy, off=(1e-7, 0.), pol='vv', size=1.):
    """One facet whose mean slope is the specular slope plus `off`."""
    Dx = np.array([[np.tan(th) * np.cos(ph) + off[0]]])
    Dy = np.array([[np.tan(th) * np.sin(ph) + off[1]]])
    E = field_spa(K0, pol, np.array([th]), np.array([ph]), Dx, Dy, np.array([[Dxx]]),
                  np.array([[Dyy]]), np.array([[Dxy]]), size, size)
    return E[0, 0], Dx[0, 0], Dy[0, 0]


@pytest.mark.parametrize('pol', ['vv', 'hh'])
@pytest.mark.parametrize('th_deg,ph,Dxx,Dyy,Dxy', [
    (30., 0., 0.2, 0.2, 0.),          # sphere cap, radius 5 m
    (20., 0.7, 1 / 3., 1 / 11., 0.),  # elongated cap
    (45., -2.0, 1.25, 0.025, 0.),
    (35., 1.2, 0.3, 0.1, 0.12),       # principal axes rotated
    (25., 0.3, 0.2, -0.05, 0.03),     # saddle (J < 0)
])
def test_specular_point_matches_geometrical_optics(th_deg, ph, Dxx, Dyy, Dxy, pol):
    th = np.deg2rad(th_deg)
    E, Dx, Dy = _facet(th, ph, Dxx, Dyy, Dxy, pol=pol)
    J = Dxx * Dyy - Dxy ** 2
    sigma_go = np.pi * (1 + Dx ** 2 + Dy ** 2) ** 2 / abs(J) * abs(R0) ** 2
    esn2 = abs(E) ** 2 * (4 * np.pi) ** 3
    assert esn2 == pytest.approx(sigma_go, rel=1e-5)


def test_stationary_point_outside_facet_gives_zero():
    th = np.deg2rad(30.)
    # the stationary point sits at delta = -(slope offset) / curvature: 0.05 / 0.2 = 0.25 m
    inside, _, _ = _facet(th, 0., 0.2, 0.2, 0., off=(0.05, 0.), size=1.)
    outside, _, _ = _facet(th, 0., 0.2, 0.2, 0., off=(0.05, 0.), size=0.4)
    assert abs(inside) > 0 and outside == 0


def test_phase_follows_stationary_point():
    """arg E changes by q . delta when the stationary point moves by delta."""
    th, ph = np.deg2rad(30.), 0.4
    c = 0.2
    d1, d2 = 1e-7, 2e-3                       # slope offsets along x
    E1, _, _ = _facet(th, ph, c, c, 0., off=(d1, 0.))
    E2, _, _ = _facet(th, ph, c, c, 0., off=(d2, 0.))
    # delta = -(offset) / c along x, delta_z = delta_x * Dx
    Dx = np.tan(th) * np.cos(ph)
    q = K0 * np.array([-2 * np.sin(th) * np.cos(ph), -2 * np.sin(th) * np.sin(ph), 2 * np.cos(th)])
    dx1, dx2 = -d1 / c, -d2 / c
    want = (q[0] * dx2 + q[2] * dx2 * (Dx + d2)) - (q[0] * dx1 + q[2] * dx1 * (Dx + d1))
    got = np.angle(E2 / E1)
    # the polarisation vector changes by O(offset) as well; the phase term dominates
    assert np.angle(np.exp(1j * (got - want))) == pytest.approx(0., abs=2e-3)
