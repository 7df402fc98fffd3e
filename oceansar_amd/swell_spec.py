"""Directional swell spectrum (Gaussian in frequency and direction).

Restates oceansar/swell_spec/dir_swell_spec.py:24-39 (ardhuin_swell_spec).
"""
import numpy as np

from . import constants as const


def ardhuin_swell_spec(k_axis, theta_axis, dir_swell_dir,
                       freq_r=0.068, sigf=0.007, sigs=8, Hs=1.45):
    f_k = 1 / 2 / np.pi * np.sqrt(const.g * k_axis)
    fac_f_k = 1 / 4 / np.pi * np.sqrt(const.g / k_axis)
    amp = (Hs / 4) ** 2 / sigf / np.sqrt(2 * np.pi)
    efs = (amp * np.exp(-(f_k - freq_r) ** 2 / (2 * sigf ** 2)) + 1E-5) * fac_f_k
    ang = np.angle(np.exp(1j * (theta_axis - np.radians(dir_swell_dir))))
    dirss = np.exp(-ang ** 2 / (2 * np.radians(sigs) ** 2))
    factor = np.sqrt(2 * np.pi) * np.radians(sigs)
    return efs * dirss / factor
