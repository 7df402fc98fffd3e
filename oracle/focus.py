"""SAR focusing of the raw cube, FP64 numpy. Test infrastructure only.

Restates the processing loop of sar_focus (oceansar/radarsim/sar_processor.py:
131-229) for one channel's planes [pol][az][rg]: zero padding to powers of
two (:131-139), range and azimuth FFTs (:167), the range-cell-migration phase
exp(4 pi i rcmc(fa) fr / c) with rcmc(fa) = sr0 / sqrt(1 - (fa lambda / 2 v)^2)
- sr0 (:165,170-171), the range IFFT (:172), the azimuth filter (weighting over
the Doppler bandwidth divided by the two-way beam, times exp(i ph_ac),
:193-213), the azimuth IFFT (:215), the crop to the original size and the
removal of the invalid azimuth samples (:219-225).
"""
import numpy as np

C = 299792458.0


def sinc_bp(sin_angle, L, f0, field=False):
    """oceansar/radarsim/antenna.py:5-25 (no squint, no beamwidth)."""
    p = np.sinc(L / (C / f0) * sin_angle)
    return p if field else p ** 2


def tables(n_az, n_rg, prf, rg_sampling, f0, v_ground, sr0, doppler_bw, az_weighting,
           ant_l_tx, ant_l_rx, ant_l=None, num_ch=1, l_total=False):
    """The 1-D tables of the loop for padded sizes (n_az, n_rg):
    fr [n_rg], fa [n_az], rcmc_fa [n_az] and the azimuth filter h [n_az] =
    exp(i ph_ac) weighting (sar_processor.py:148-213). ant_l: the cfg's ant_L
    (sinc_1tx_nrx branch, :156-160, whose `field` argument lands in sinc_bp's
    *dummy, so it returns the power pattern), else the ant_L_tx/ant_L_rx branch."""
    l0 = C / f0
    fr = np.fft.fftfreq(n_rg, 1 / rg_sampling)
    fa = np.fft.fftfreq(n_az, 1 / prf)
    sin_az = fa * l0 / (2 * v_ground)
    if ant_l is not None:
        L = ant_l * num_ch if l_total else ant_l
        n = num_ch if l_total else 1
        beam = sinc_bp(sin_az, L, f0) * sinc_bp(sin_az, L / n, f0)
    else:
        beam = sinc_bp(sin_az, ant_l_tx, f0, field=True) * sinc_bp(sin_az, ant_l_rx, f0, field=True)
    rcmc_fa = sr0 / np.sqrt(1 - (fa * (l0 / 2.) / v_ground) ** 2.) - sr0
    n_samp = 2 * (int(doppler_bw / (fa[1] - fa[0])) / 2)
    weighting = (az_weighting -
                 (1. - az_weighting) * np.cos(2 * np.pi * np.linspace(0, 1., int(n_samp))))
    L_win = np.sum(np.abs(weighting) ** 2) / weighting.size
    weighting /= np.sqrt(L_win)
    if fa.size > n_samp:
        zeros = np.zeros(n_az)
        zeros[0:int(n_samp)] = weighting
        weighting = np.roll(zeros, int(-n_samp / 2))
    with np.errstate(divide='ignore', invalid='ignore'):
        weighting = np.where(np.abs(beam) > 0, weighting / beam, 0)
    ph_ac = 4. * np.pi / l0 * sr0 * (np.sqrt(1. - (fa * l0 / 2. / v_ground) ** 2.) - 1.)
    h = np.exp(1j * ph_ac) * weighting
    return fr, fa, rcmc_fa, h


def valid_rows(az_size_orig, doppler_bw, v_ground, f0, sr0, prf):
    """[lo, hi) of the azimuth samples kept after focusing (sar_processor.py:222-225)."""
    l0 = C / f0
    n_val_az_2 = np.floor(doppler_bw / 2. / (2. * v_ground ** 2. / l0 / sr0) * prf / 2.) * 2.
    return int(n_val_az_2), int(az_size_orig - n_val_az_2 - 1)


def focus_channel(raw_ch, prf, rg_sampling, f0, v_ground, sr0, doppler_bw, az_weighting,
                  ant_l_tx, ant_l_rx, ant_l=None, num_ch=1, l_total=False):
    """SLC of one channel: raw_ch [pol][az][rg] (complex) -> [pol][az_valid][rg]."""
    raw_ch = np.asarray(raw_ch)
    npol, az0, rg0 = raw_ch.shape
    # optimize_fftsize with max_prime 2 (utils/misc.py:56-74): 2**ceil(log2 n)
    n_az, n_rg = (2 ** np.ceil(np.log2([az0, rg0]))).astype(int)
    data = np.zeros((npol, n_az, n_rg), dtype=complex)
    data[:, :az0, :rg0] = raw_ch
    fr, fa, rcmc_fa, h = tables(n_az, n_rg, prf, rg_sampling, f0, v_ground, sr0, doppler_bw,
                                az_weighting, ant_l_tx, ant_l_rx, ant_l, num_ch, l_total)
    data = np.fft.fft(np.fft.fft(data, axis=-1), axis=-2)
    data = data * np.exp(4j * np.pi * rcmc_fa.reshape((1, n_az, 1)) / C * fr.reshape((1, 1, n_rg)))
    data = np.fft.ifft(data, axis=2)
    data = data * h.reshape((1, n_az, 1))
    data = np.fft.ifft(data, axis=1)
    data = data[:, :az0, :rg0]
    lo, hi = valid_rows(az0, doppler_bw, v_ground, f0, sr0, prf)
    return data[:, lo:hi, :]
