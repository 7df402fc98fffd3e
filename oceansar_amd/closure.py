"""Sub-grid closure: coherence time and random-scatterer time series.

grid_coherence restates oceansar/closure/coherence_times.py:7-18.
RandomScatterers restates oceansar/closure/random_scatterers.py:6-31 with
explicit, seeded random inputs instead of numpy's global RNG (the reference
never seeds it) and an integer pulse index in place of int(t * Fs)
(random_scatterers.py:29 can truncate t*Fs to pulse-1).
"""
import numpy as np
from scipy import special

from . import constants as const
from .utils import optimize_fftsize


def grid_coherence(U, res, f0):
    """Pierson-Moskowitz coherence time of a facet (coherence_times.py:7-18)."""
    return 3.29 * const.c / f0 / U / np.sqrt(special.erf(2.688 * res / U ** 2))


class RandomScatterers:
    """Unit-power complex Gaussian-spectrum series s[N] and start offsets ind0.

    Parameters
    ----------
    tau : coherence time [s]; dim : (Ny, Nx); Fs : PRF [Hz]
    rng : numpy Generator supplying, in this order, the 2N normals of the
          spectrum and the Ny*Nx uniforms of the start offsets.
    """

    def __init__(self, tau, dim, Fs, rng, N=None):
        if N is None:
            N = int(tau * Fs * 200)
        self.N_unopt = int(N)
        self.N = int(optimize_fftsize(N, max_prime=5))
        f = np.fft.fftfreq(self.N, 1. / Fs)
        a = 1 / (tau ** 2)
        spec = np.sqrt(np.pi / a) * np.exp(-np.pi ** 2 * f ** 2 / a)
        s_f = rng.standard_normal(self.N) + 1j * rng.standard_normal(self.N)
        s = np.fft.ifft(s_f * np.sqrt(spec))
        self.s = s / np.sqrt(np.mean(np.abs(s) ** 2))
        self.ind0 = (self.N_unopt * rng.uniform(size=dim)).astype(np.int64)
        self.Fs = Fs

    def scats_at_pulse(self, pulse):
        """s[(ind0 + pulse) mod N] for an integer pulse index."""
        return self.s[np.mod(self.ind0 + int(pulse), self.N)]
