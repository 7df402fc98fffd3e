"""Range impulse response table and the ``chan_profile`` entry point.

calc_sinc_vec restates oceansar/radarsim/range_profile.py:15-19: an
n_sinc_samples-tap sinc, oversampled sinc_ovs times, as float32. The table is
an input of the echo accumulator; the GPU uses exactly these float32 values.

chan_profile mirrors chan_profile_numba (range_profile.py:83-116): it adds
the 8-tap responses of all facets of one (pulse, channel) into ``output`` in
place, here on the device through the HIP echo-deposit kernel.
"""
import numpy as np

N_SINC_SAMPLES = 8
SINC_OVS = 20


def calc_sinc_vec(n_sinc_samples=6, sinc_ovs=10, Fs=1.0):
    n_bins = int(sinc_ovs * n_sinc_samples + 1)
    bins = np.arange(n_bins) * 1. / sinc_ovs - n_sinc_samples / 2
    return np.array(np.sinc(bins / Fs), dtype=np.float32)


def chan_profile(srg, scene, binsize, min_sr, sinc_vec, n_sinc_samples,
                 sinc_ovs, output, rg_only=False):
    """Device drop-in for chan_profile_numba (range_profile.py:83-116).

    srg: float64 device tensor (any shape, flattened); scene: complex64 or
    complex128 device tensor of the same size; output: complex64 device
    tensor, accumulated in place. Raises TypeError on dtype/shape mismatch,
    like the numba signature does.
    """
    from . import _lib
    return _lib.echo_deposit(srg, scene, binsize, min_sr, sinc_vec,
                             n_sinc_samples, sinc_ovs, output, rg_only)
