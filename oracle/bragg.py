"""Bragg RCS of every facet, Romeiser & Alpers (1997) composite model, FP64.

Restates RCSRomeiser97.rcs_numba / _rcs_numba
(oceansar/nrcs/romeiser97.py:251-393): tilted slopes, local incidence,
complex b_hh / b_vv, weighting w, transfer functions T, Bragg wave number
and direction, linear interpolation of the spectrum table with fill value 1
(romeiser97.py:60-65, scipy interp1d semantics), spreading, bad-facet mask.
"""
import numpy as np

EPS_SW = 73 + 18j
G = 9.80665
_rho, _S, _X_0 = 1000., 0.072, 22e3


def spread_elfouhaily(k, theta, U_10, fetch):
    """oceansar/spread/elfouhaily.py:23-56."""
    k_0 = G / U_10 ** 2
    X = k_0 * fetch
    Omega_c = 0.84 * np.tanh((X / _X_0) ** 0.4) ** (-0.75)
    with np.errstate(divide='ignore', invalid='ignore'):
        cK = np.sqrt(G / k + _S / _rho * k)
        k_p = k_0 * Omega_c ** 2
        cK_p = np.sqrt(G / k_p + _S / _rho * k_p)
        k_m = np.sqrt(_rho * G / _S)
        cK_m = np.sqrt(G / k_m + _S / _rho * k_m)
        C_10 = (0.8 + 0.065 * U_10) * 1e-3
        ustar = np.sqrt(C_10) * U_10
        a_0 = np.log(2.) / 4.
        a_m = 0.13 * ustar / cK_m
        Delta = np.tanh(a_0 + 4. * (cK / cK_p) ** 2.5 + a_m * (cK_m / cK) ** 2.5)
    return np.where((theta > -np.pi / 2.) & (theta < np.pi / 2.),
                    (1. + Delta * np.cos(2. * theta)) / np.pi, 0)


def spread_cos2s(k, theta, U_10, fetch):
    """oceansar/spread/cos2s.py:13-18 with an elementwise mask."""
    G_ = 1. / np.pi * np.cos(theta) ** 2
    return np.where((theta < -np.pi / 2.) | (theta > np.pi / 2.), 0., G_)


SPREADS = {0: spread_elfouhaily, 1: spread_cos2s}


def interp_fill1(x, xp, fp):
    """scipy interp1d(kind='linear', bounds_error=False, fill_value=1)."""
    x = np.asarray(x, dtype=np.float64)
    idx = np.searchsorted(xp, x, side='left')
    idx = np.clip(idx, 1, len(xp) - 1)
    lo, hi = idx - 1, idx
    slope = (fp[hi] - fp[lo]) / (xp[hi] - xp[lo])
    y = slope * (x - xp[lo]) + fp[lo]
    out = (x < xp[0]) | (x > xp[-1])
    return np.where(out, 1.0, y)


def facet_terms(inc, az, diffx, diffy, k0, wind_dir, d):
    """_rcs_numba body (romeiser97.py:262-314): k_b, theta_inp[2], T_hh, T_vv, bad."""
    cos_az = np.cos(az)
    sin_az = np.sin(az)
    s_p = np.arctan(cos_az * diffx + sin_az * diffy)
    s_t = np.arctan(-sin_az * diffx + cos_az * diffy)
    cos_s_t = np.cos(s_t)
    cos_s_t_2 = cos_s_t ** 2
    sin_s_t_2 = 1 - cos_s_t_2
    sin_ip = np.sin(inc - s_p)
    cos_ip = np.cos(inc - s_p)
    cos_ip_2 = cos_ip ** 2
    sin_ip_2 = sin_ip ** 2
    cos_ti = cos_ip * cos_s_t
    sin_ti_2 = 1 - cos_ti ** 2
    sq = np.sqrt(EPS_SW)
    b_hh = EPS_SW / (cos_ti + sq) ** 2.
    b_vv = ((EPS_SW ** 2.) * (1. + sin_ti_2)) / ((EPS_SW * cos_ti + sq) ** 2.)
    w = cos_ip / (np.cos(inc) * np.cos(s_p))
    a = sin_s_t_2 / sin_ti_2
    b = sin_ip_2 * cos_s_t_2 / sin_ti_2
    F0 = 8 * np.pi * (k0 * cos_ti) ** 4
    T_vv = w * (F0 * np.abs(b * b_vv + a * b_hh) ** 2.)
    T_hh = w * (F0 * np.abs(b * b_hh + a * b_vv) ** 2.)
    k_b = 2 * k0 * np.sqrt(sin_ip_2 + cos_ip_2 * sin_s_t_2)
    k_b = np.where(k_b > 2 * k0, 0, k_b)
    phi_b = az + np.arctan((cos_ip * np.sin(-s_t)) / np.sin(inc - s_p))
    th = np.empty((2,) + phi_b.shape)
    th[0] = phi_b - wind_dir
    th[1] = phi_b + np.pi - wind_dir
    th = np.angle(np.exp(1j * th))
    bad = 4 * sin_ip_2 < d ** 2
    return k_b, th, T_hh, T_vv, bad


def rcs_romeiser97(inc_col, az_row, diffx, diffy, k0, pol, wind_dir, d, spec_k, spec_S,
                   spread_id, wind_U, wind_fetch, dx, dy):
    """Bragg RCS[2, Ny, Nx] per polarisation as a dict {'hh': ..., 'vv': ...}.

    inc_col [Nx] and az_row [Ny] are broadcast (romeiser97.py:331-339);
    index 0 is the minus component (theta = phi_b - wind_dir).
    """
    inc = np.asarray(inc_col, dtype=np.float64)[None, :]
    az = np.asarray(az_row, dtype=np.float64)[:, None]
    diffx = np.asarray(diffx, dtype=np.float64)
    diffy = np.asarray(diffy, dtype=np.float64)
    k_b, th, T_hh, T_vv, bad = facet_terms(inc, az, diffx, diffy, k0, wind_dir, d)
    return rcs_from_terms(k_b, th, T_hh, T_vv, bad, pol, spec_k, spec_S, spread_id, wind_U,
                          wind_fetch, dx, dy)


def rcs_from_terms(k_b, th, T_hh, T_vv, bad, pol, spec_k, spec_S, spread_id, wind_U,
                   wind_fetch, dx, dy):
    """RCSRomeiser97.rcs_numba after _rcs_numba (romeiser97.py:338-393): folded
    spectrum S(k_b)/k_b G(k_b, theta), RCS = T spectrum dx dy, bad-facet mask.
    k_b = 0 (the reference's k_b > 2 k0 rule) gives S = 1 (fill value) over 0:
    inf where G > 0, NaN where G = 0, before the mask."""
    k_inp = np.array([k_b, k_b])
    S1 = interp_fill1(k_b.ravel(), spec_k, spec_S).reshape((1,) + k_b.shape)
    with np.errstate(divide='ignore', invalid='ignore'):
        spectrum = S1 / k_inp * SPREADS[spread_id](k_inp, th, wind_U, wind_fetch)
    out = {}
    with np.errstate(invalid='ignore'):
        if pol in ('vv', 'DP'):
            r = T_vv * spectrum * np.float32(dx) * np.float32(dy)
            r[0][bad] = 0
            r[1][bad] = 0
            out['vv'] = r
        if pol in ('hh', 'DP'):
            r = T_hh * spectrum * np.float32(dx) * np.float32(dy)
            if pol == 'DP':
                # romeiser97.py:391-392: only the plus component is masked for DP
                r[1][bad] = 0
            else:
                r[0][bad] = 0
                r[1][bad] = 0
            out['hh'] = r
    return out
