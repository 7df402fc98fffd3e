"""Spherical-Earth SAR geometry on the host, in FP64.

Restates oceansar/utils/geometry.py:9-173. gr_to_geo keeps the reference's
500-point look-angle table and scipy's quadratic InterpolatedUnivariateSpline,
because the slant range it returns is the FP64 range reference of every facet
and must match the reference to the last few ulps.
"""
import numpy as np
from scipy import interpolate

from . import constants as const


def orbit_to_vel(orbit_alt, ground=False, r_planet=const.r_earth,
                 m_planet=const.m_earth):
    """Circular-orbit velocity (geometry.py:9-25)."""
    v = np.sqrt(const.G * m_planet / (r_planet + orbit_alt))
    if ground:
        v = r_planet / (r_planet + orbit_alt) * v
    return v


def inc_to_look(theta_i, orbit_alt, r_planet=const.r_earth):
    """Look angle from incidence angle (geometry.py:60-70)."""
    return np.arcsin(np.sin(theta_i) / (r_planet + orbit_alt) * r_planet)


def look_to_inc(theta_l, orbit_alt, r_planet=const.r_earth):
    """Incidence angle from look angle (geometry.py:73-82)."""
    return np.arcsin(np.sin(theta_l) * (r_planet + orbit_alt) / r_planet)


def inc_to_sr(theta_i, orbit_alt, r_planet=const.r_earth):
    """Slant range from incidence angle (geometry.py:28-43)."""
    theta_l = inc_to_look(theta_i, orbit_alt, r_planet=r_planet)
    dth = theta_i - theta_l
    return np.sqrt((orbit_alt + r_planet - r_planet * np.cos(dth)) ** 2 +
                   (r_planet * np.sin(dth)) ** 2)


def inc_to_gr(theta_i, orbit_alt, r_planet=const.r_earth):
    """Ground range from incidence angle (geometry.py:46-57)."""
    return r_planet * (theta_i - inc_to_look(theta_i, orbit_alt,
                                             r_planet=r_planet))


def max_look_angle(orbit_alt, r_planet=const.r_earth):
    """Horizon look angle (geometry.py:165-173)."""
    return np.arcsin(r_planet / (r_planet + orbit_alt))


def gr_to_geo(ground_range, orbit_alt, r_planet=const.r_earth):
    """Slant range, incidence and look angle at ground ranges (geometry.py:136-162).

    Note the reference uses const.r_earth (not r_planet) in the first term of
    the slant-range expression; kept as is.
    """
    theta_l = np.linspace(0, max_look_angle(orbit_alt), 500)
    theta_i = look_to_inc(theta_l, orbit_alt, r_planet=r_planet)
    dth = theta_i - theta_l
    gr = r_planet * dth
    sr = np.sqrt((orbit_alt + const.r_earth - r_planet * np.cos(dth)) ** 2 +
                 (r_planet * np.sin(dth)) ** 2)
    spl = interpolate.InterpolatedUnivariateSpline
    theta_l_i = spl(gr, theta_l, k=2)(ground_range)
    theta_i_i = spl(gr, theta_i, k=2)(ground_range)
    sr_i = spl(gr, sr, k=2)(ground_range)
    return sr_i, theta_i_i, theta_l_i


def sr_to_geo(slant_range, orbit_alt, r_planet=const.r_earth, m_planet=const.m_earth):
    """Ground range, incidence angle, look angle and the centripetal term
    v_orb^2 r_track / (r_planet + h) at slant ranges (geometry.py:101-133):
    a 2500-point look-angle table from nadir to the horizon and scipy's
    quadratic InterpolatedUnivariateSpline over slant range, as the reference."""
    theta_l = np.linspace(0, max_look_angle(orbit_alt, r_planet=r_planet), 2500)
    theta_i = look_to_inc(theta_l, orbit_alt, r_planet=r_planet)
    dth = theta_i - theta_l
    r_track = np.cos(dth) * r_planet
    v_orb = orbit_to_vel(orbit_alt, r_planet=r_planet, m_planet=m_planet)
    b = v_orb ** 2 * r_track / (r_planet + orbit_alt)
    gr = r_planet * dth
    sr = np.sqrt((orbit_alt + r_planet - r_planet * np.cos(dth)) ** 2 +
                 (r_planet * np.sin(dth)) ** 2)
    spl = interpolate.InterpolatedUnivariateSpline
    return (spl(sr, gr, k=2)(slant_range), spl(sr, theta_i, k=2)(slant_range),
            spl(sr, theta_l, k=2)(slant_range), spl(sr, b, k=2)(slant_range))
