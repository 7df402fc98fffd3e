"""Directional spreading functions, registry ``models[name](k, theta, U_10, fetch)``.

Mirrors oceansar/spread/__init__.py:51. On the path: Elfouhaily
(oceansar/spread/elfouhaily.py:23-56) and cos^2 (oceansar/spread/cos2s.py:13-18).
The same formulas run per facet inside the Bragg facet kernel
(oceansar_amd/csrc/bragg.hip); this host version is used when the
surface spectrum is built.
"""
import numpy as np

from .constants import g

_rho = 1000.
_S = 0.072
_X_0 = 22e3


def elfouhaily_delta_scalars(U_10, fetch):
    """The wave-number independent terms of Delta(k) (elfouhaily.py:26-52):
    a_0, a_p, a_m, cK_p, cK_m."""
    k_0 = g / U_10 ** 2
    X = k_0 * fetch
    Omega_c = 0.84 * np.tanh((X / _X_0) ** 0.4) ** (-0.75)
    k_p = k_0 * Omega_c ** 2
    cK_p = np.sqrt(g / k_p + _S / _rho * k_p)
    k_m = np.sqrt(_rho * g / _S)
    cK_m = np.sqrt(g / k_m + _S / _rho * k_m)
    C_10 = (0.8 + 0.065 * U_10) * 1e-3
    ustar = np.sqrt(C_10) * U_10
    return dict(a_0=np.log(2.) / 4., a_p=4., a_m=0.13 * ustar / cK_m, cK_p=cK_p, cK_m=cK_m)


def elfouhaily_delta(k, U_10, fetch):
    """Delta(k) of the Elfouhaily spreading, eq. 57 (elfouhaily.py:26-52)."""
    s = elfouhaily_delta_scalars(U_10, fetch)
    cK = np.sqrt(g / k + _S / _rho * k)
    return np.tanh(s['a_0'] + s['a_p'] * (cK / s['cK_p']) ** 2.5 +
                   s['a_m'] * (s['cK_m'] / cK) ** 2.5)


def elfouhaily(k, theta, U_10, fetch):
    """G = (1 + Delta cos 2theta)/pi on the upwind half plane, else 0 (elfouhaily.py:54)."""
    k = np.asarray(k, dtype=np.float64)
    theta = np.asarray(theta, dtype=np.float64)
    delta = elfouhaily_delta(k, float(U_10), float(fetch))
    return np.where((theta > -np.pi / 2.) & (theta < np.pi / 2.),
                    (1. + delta * np.cos(2. * theta)) / np.pi, 0.)


def cos2s(k, theta, U_10, fetch):
    """G = cos^2(theta)/pi for |theta| <= pi/2, else 0 (cos2s.py:13-18).

    The reference masks with Python ``or`` on arrays, which raises; the mask
    is applied elementwise here, keeping the reference's 1/2 normalisation.
    """
    theta = np.asarray(theta)
    G = 1. / np.pi * np.cos(theta) ** 2
    G = np.where((theta < -np.pi / 2.) | (theta > np.pi / 2.), 0., G)
    return G


models = {'elfouhaily': elfouhaily, 'cos2s': cos2s, 'none': None}
SPREAD_IDS = {'elfouhaily': 0, 'cos2s': 1}
