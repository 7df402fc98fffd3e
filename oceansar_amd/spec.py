"""Omnidirectional wave-number spectra, registry ``models[name](k, U_10, fetch)``.

Mirrors the registry shape of oceansar/spec/__init__.py:38. Only the
Elfouhaily et al. (1997) spectrum is on the raw-data path
(oceansar/spec/elfouhaily.py:43-99); it is evaluated on the host, once per
scene, for the surface coefficients and for the Bragg spectrum table.
Arithmetic follows numpy's dtype rules exactly as the reference does
(float32 wave numbers give a float32 spectrum).
"""
import warnings

import numpy as np

from .constants import g

_rho = 1000.     # water density, elfouhaily.py:15
_S = 0.072       # surface tension, elfouhaily.py:16
_X_0 = 22e3      # dimensionless fetch constant, elfouhaily.py:17


def _gamma(Omega_c):
    # elfouhaily.py:20-30 (peak enhancement; out-of-range uses Omega_c = 5)
    if 0.84 < Omega_c < 1.0:
        return 1.7
    if 1.0 < Omega_c < 5.0:
        return 1.7 + 6. * np.log10(Omega_c)
    warnings.warn('Omega_c is out of range. Returning  value for 5.0',
                  RuntimeWarning)
    return 1.7 + 6. * np.log10(5.0)


def _alpha_m(ustar, c_m):
    # elfouhaily.py:33-40
    if ustar < c_m:
        return 1e-2 * (1. + np.log(ustar / c_m))
    return 1e-2 * (1. + 3. * np.log(ustar / c_m))


def elfouhaily_scalars(U_10, fetch):
    """The wave-number independent terms of the Elfouhaily spectrum
    (elfouhaily.py:43-99), with the reference's scalar types: k_m, k_p, c_p,
    Omega, alpha_p, sigma, gamma, alpha_m, c_m."""
    k_m = 2 * np.pi / 0.017
    k_0 = g / U_10 ** 2
    X = k_0 * fetch
    Omega_c = 0.84 * np.tanh((X / _X_0) ** 0.4) ** (-0.75)
    k_p = k_0 * Omega_c ** 2
    c_p = np.sqrt(g / k_p + _S / _rho * k_p)
    Omega = U_10 / c_p
    alpha_p = 6e-3 * np.sqrt(Omega)
    sigma = 0.08 * (1. + 4. * Omega_c ** (-3.))
    C_10 = (0.8 + 0.065 * U_10) * 1e-3
    ustar = np.sqrt(C_10) * U_10
    c_m = np.sqrt(g / k_m + _S / _rho * k_m)
    return dict(k_m=k_m, k_p=k_p, c_p=c_p, Omega=Omega, alpha_p=alpha_p, sigma=sigma,
                gamma=_gamma(Omega_c), alpha_m=_alpha_m(ustar, c_m), c_m=c_m)


def elfouhaily(k, U_10, fetch):
    """Elfouhaily omnidirectional spectrum S(k) = (B_l + B_h)/k^3 (elfouhaily.py:43-99)."""
    s = elfouhaily_scalars(U_10, fetch)
    k_m, k_p, c_p, Omega = s['k_m'], s['k_p'], s['c_p'], s['Omega']
    alpha_p, sigma, alpha_m, c_m = s['alpha_p'], s['sigma'], s['alpha_m'], s['c_m']
    c = np.sqrt(g / k + _S / _rho * k)
    Gamma = np.exp(-(np.sqrt(k / k_p) - 1.) ** 2 / (2. * sigma ** 2))
    J_p = s['gamma'] ** Gamma
    L_pm = np.exp(-5. / 4. * (k_p / k) ** 2)
    F_p = L_pm * J_p * np.exp(-Omega / np.sqrt(10.) * (np.sqrt(k / k_p) - 1.))
    B_l = 0.5 * alpha_p * c_p / c * F_p
    F_m = L_pm * np.exp(-0.25 * (k / k_m - 1.) ** 2)
    B_h = 0.5 * alpha_m * c_m / c * F_m
    return (B_l + B_h) / k ** 3


models = {'elfouhaily': elfouhaily}
