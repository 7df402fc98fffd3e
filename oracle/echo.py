"""Range-response table and the per-facet 8-tap integrator.

calc_sinc_vec: oceansar/radarsim/range_profile.py:15-19.
profile_integrator: range_profile.py:22-49 (numba, serial). Restated here as
eight vectorised scatter-adds, one per tap; taps outside [0, nbins) dropped.
chan_profile: range_profile.py:83-116 (FP64 bin position, floor, sub-bin
round-half-to-even).
"""
import numpy as np


def calc_sinc_vec(n_sinc_samples=8, sinc_ovs=20, Fs=1.0):
    n_bins = int(sinc_ovs * n_sinc_samples + 1)
    bins = np.arange(n_bins) * 1. / sinc_ovs - n_sinc_samples / 2
    return np.array(np.sinc(bins / Fs), dtype=np.float32)


def bin_index(srg, binsize, min_sr, sinc_ovs=20):
    """(integer bin, sub-bin) of every facet (range_profile.py:106-108)."""
    f = (np.asarray(srg, dtype=np.float64) - min_sr) / float(binsize)
    d = np.int32(np.floor(f))
    sb = np.int32(np.round(sinc_ovs * (1 - (f - d))))
    return d, sb


def profile_integrator(scat, bin_i, bin_f, pulse, n_samp, over_samp, output):
    """output[bin_i - n_samp//2 + j] += pulse[bin_f + j*over_samp] * scat, j < n_samp."""
    nbins = output.size
    scat = np.asarray(scat, dtype=np.complex128)
    base = np.asarray(bin_i, dtype=np.int64) - int(n_samp / 2)
    sb = np.asarray(bin_f, dtype=np.int64)
    pulse = np.asarray(pulse, dtype=np.float32)
    acc = np.zeros(nbins, dtype=np.complex128)
    for j in range(n_samp):
        b = base + j
        ok = (b >= 0) & (b < nbins)
        w = pulse[sb[ok] + j * over_samp].astype(np.float64)
        np.add.at(acc, b[ok], w * scat[ok])
    output += acc


def chan_profile(srg, scene, binsize, min_sr, sinc_vec, n_sinc_samples, sinc_ovs, output):
    d, sb = bin_index(np.ravel(srg), binsize, min_sr, sinc_ovs)
    profile_integrator(np.ravel(scene), d, sb, sinc_vec, n_sinc_samples, sinc_ovs, output)
