"""Kirchhoff approximation, stationary-phase (spa) monostatic field, FP64.

Restates RCSKA.field (oceansar/nrcs/ka.py:105-284) for the monostatic call of
sar_raw (sar_raw_nompi.py:403-426: R_i = R_s = 1, theta_s = theta_i, the
phi_s argument is ignored). np.complex (ka.py:258, removed from numpy) is
complex128 here.
"""
import numpy as np

EPS_SW = 73 + 18j


def _cross(a, b):
    return np.stack([a[1] * b[2] - a[2] * b[1],
                     a[2] * b[0] - a[0] * b[2],
                     a[0] * b[1] - a[1] * b[0]])


def field_spa(k0, pol, theta_i, phi_i, Diffx, Diffy, Diffxx, Diffyy, Diffxy, dx, dy):
    """E[Ny, Nx] (complex128), non-zero on the stationary-phase facets.

    theta_i: [Nx] incidence per column; phi_i: [Ny] azimuth per row;
    pol: 'hh' or 'vv' (pol_i == pol_s).
    """
    ny, nx = Diffx.shape
    th = np.asarray(theta_i, dtype=np.float64).reshape(1, nx)
    ph = np.asarray(phi_i, dtype=np.float64).reshape(ny, 1)
    Dx = np.asarray(Diffx, np.float64)
    Dy = np.asarray(Diffy, np.float64)
    Dxx = np.asarray(Diffxx, np.float64)
    Dyy = np.asarray(Diffyy, np.float64)
    Dxy = np.asarray(Diffxy, np.float64)
    sin_t, cos_t, tan_t = np.sin(th), np.cos(th), np.tan(th)
    sin_p, cos_p = np.sin(ph), np.cos(ph)
    shp = (ny, nx)
    one = np.ones(shp)
    h_i = np.stack([-sin_p * one, cos_p * one, 0. * one])
    v_i = np.stack([-cos_t * cos_p, -cos_t * sin_p, -sin_t * one])
    a_i = h_i if pol[0] == 'h' else v_i
    a_s = h_i if pol[1] == 'h' else v_i
    n_norm = np.sqrt(Dx ** 2. + Dy ** 2. + 1.)
    n = np.stack([-Dx / n_norm, -Dy / n_norm, 1. / n_norm])
    n_i = np.stack([sin_t * cos_p, sin_t * sin_p, -cos_t * one])
    n_s = np.stack([-sin_t * cos_p, -sin_t * sin_p, cos_t * one])
    q = k0 * (n_s - n_i)
    t = _cross(n_i, n)
    t = t / np.sqrt(np.sum(t ** 2, axis=0))
    d = _cross(n_i, t)
    cos_l = -np.sum(n * n_i, axis=0)
    sqrt_e_sin = np.sqrt(EPS_SW + cos_l ** 2. - 1.)
    r_v = (EPS_SW * cos_l - sqrt_e_sin) / (EPS_SW * cos_l + sqrt_e_sin)
    r_h = (cos_l - sqrt_e_sin) / (cos_l + sqrt_e_sin)
    a_t = np.sum(a_i * t, axis=0)
    a_d = np.sum(a_i * d, axis=0)
    n_p_ni = -cos_l
    n_x_t = _cross(n, t)
    n_x_Es = ((1. + r_h) * a_t) * n_x_t - ((1. - r_v) * n_p_ni * a_d) * t
    eta_n_x_Hs = -(((1. - r_h) * n_p_ni * a_t) * t + ((1. + r_v) * a_d) * n_x_t)
    p = np.sum(a_s * _cross(n_s, n_x_Es - _cross(n_s, eta_n_x_Hs)), axis=0)
    K = -1j * k0 * np.exp(-1j * k0 * 2.) / ((4. * np.pi) ** 2.)
    J = Dxx * Dyy - Dxy ** 2
    J = np.where(J == 0., np.nan, J)
    J_abs = np.abs(J)
    with np.errstate(invalid='ignore', divide='ignore'):
        delta_x = (1. / J_abs) * (Dxy * (Dy - tan_t * sin_p) - Dyy * (Dx - tan_t * cos_p))
        delta_y = (1. / J_abs) * (Dxy * (Dx - tan_t * cos_p) - Dxx * (Dy - tan_t * sin_p))
    delta_z = delta_x * Dx + delta_y * Dy
    epsilon = np.where(J > 0., np.sign(Dxx), 1j)
    E = np.zeros(shp, dtype=np.complex128)
    sps = np.where((-0.5 * dx < delta_x) & (delta_x < 0.5 * dx) &
                   (-0.5 * dy < delta_y) & (delta_y < 0.5 * dy))
    if sps[0].size > 0:
        E[sps] = (epsilon[sps] * p[sps] * K *
                  np.exp(1j * (q[0][sps] * delta_x[sps] + q[1][sps] * delta_y[sps] +
                               q[2][sps] * delta_z[sps])) *
                  2. * np.pi / q[2][sps] * n_norm[sps] / np.sqrt(J_abs[sps]))
    return E
