"""Uniform-aperture antenna pattern (oceansar/radarsim/antenna.py:5-26)."""
import numpy as np
from scipy.constants import c


def sinc_bp(sin_angle, L, f0, field=False, beamwidth=None, sin_squint=0):
    """sinc(L/lambda * (sin_angle - sin_squint)); squared unless field=True."""
    lam = c / f0
    La = L if beamwidth is None else lam / beamwidth
    pattern = np.sinc(La / lam * (sin_angle - sin_squint))
    return pattern if field else pattern ** 2
