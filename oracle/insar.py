"""InSAR L1b and ATI processing of focused SLC data, FP64 numpy. Test
infrastructure only.

Restates, without plots:
  smooth1d / smooth      oceansar/utils/misc.py:108-197 (flat windows: the
                         edge-normalised moving average summed in offset order;
                         other windows: scipy fftconvolve 'same' with the
                         normalised numpy window, zero beyond the edges)
  uwphase                oceansar/radarsim/insar_processor.py:31-39 (also
                         ati_processor.py:31-39)
  crop                   insar_processor.py:119-163 (region of interest)
  coregister             insar_processor.py:165-174 (sub-sample azimuth shift
                         of every channel through the FFT; the shifted values
                         are stored back into the complex64 array read from
                         the SLC file, as the reference's in-place assignment
                         does; the FFTs are complex128 as numpy < 1.24 computes
                         them for complex64 input)
  flattening_phasor      insar_processor.py:176-185
  l1b                    insar_processor.py:186-217 (intensities, pairwise
                         interferograms, coherences and the coherence LUT)
  ati_stats              oceansar/radarsim/ati_processor.py:97-470 (the
                         velocity and NRCS statistics that ati_process writes)
"""
import numpy as np
from scipy import signal

C = 299792458.0
WINDOWS = ('flat', 'hanning', 'hamming', 'bartlett', 'blackman')


def smooth1d(data, window_len, axis=0):
    """misc.py:108-143: out[n] = sum of data[n + d] for d = -(wl // 2) ..
    wl - wl // 2 - 1 inside the array, in that order, divided by the count."""
    shp = np.array(data.shape).astype(int)
    out = np.zeros(shp, dtype=data.dtype)
    wl = int(window_len)
    wlh1 = int(window_len / 2)
    normf = np.zeros(shp[axis])
    n = shp[axis]
    for ind in range(wl):
        d = ind - wlh1
        lo, hi = max(0, -d), min(n, n - d)      # outputs with n + d inside [0, n)
        if hi <= lo:
            continue
        normf[lo:hi] += 1
        src = [slice(None)] * data.ndim
        dst = [slice(None)] * data.ndim
        src[axis] = slice(lo + d, hi + d)
        dst[axis] = slice(lo, hi)
        out[tuple(dst)] += data[tuple(src)]
    shpn = np.ones_like(shp)
    shpn[axis] = shp[axis]
    out /= normf.reshape(shpn)
    return out


def window_taps(window, window_len):
    """Normalised taps of a non-flat window (misc.py:181-195)."""
    w = getattr(np, window)(int(window_len))
    return w / w.sum()


def smooth(data, window_len_=11, window='flat', axis=None):
    """misc.py:148-197 for the calls of the processors (axis given)."""
    window_len = int(window_len_)
    if data.ndim > 2:
        raise ValueError('Arrays with ndim > 2 not supported')
    if window not in WINDOWS:
        raise ValueError('Window type not supported')
    if window == 'flat' and (axis is not None or data.ndim == 1):
        return smooth1d(data, window_len, axis=0 if axis is None else axis)
    if window_len < 3:
        return data
    w = np.ones(window_len) if window == 'flat' else getattr(np, window)(window_len)
    if data.ndim > 1:
        if axis is None:
            w = np.sqrt(np.outer(w, w))
        elif axis == 0:
            w = w.reshape((w.size, 1))
        else:
            w = w.reshape((1, w.size))
    return signal.fftconvolve(data, w / w.sum(), mode='same')


def uwphase(phasor):
    """insar_processor.py:31-39: phases re-centred on the left edge of the
    fullest histogram bin (the reference averages an edge with itself)."""
    pha = np.angle(phasor)
    nbins = np.min([72, np.ceil(phasor.size / 50).astype(int)])
    pha_hist, h_edgs = np.histogram(pha, bins=nbins, range=(-np.pi, np.pi))
    pha_mod = (h_edgs[pha_hist.argmax()] + h_edgs[pha_hist.argmax()]) / 2
    return np.angle(phasor * np.exp(-1j * pha_mod)) + pha_mod


def radial_velocity(vx, vz, inc_angle):
    """insar_processor.py:126 / ati_processor.py:129 under the reference's numpy
    (< 1.24, value-based casting): float32 fields times float64 scalars are
    float32 products with the scalars rounded to float32."""
    return vx * np.float32(np.sin(inc_angle)) - vz * np.float32(np.cos(inc_angle))


def velocity_stats(vx, vz, inc_angle, v_ground, sr0):
    """insar_processor.py:126-142: the radial surface velocity, its mean and
    standard deviation (numpy float32 reductions of the float32 fields), and
    the expected azimuth shift and its spread."""
    v_radial = radial_velocity(vx, vz, inc_angle)
    mean, std = np.mean(v_radial), np.std(v_radial)
    return v_radial, mean, std, -mean / v_ground * sr0, std / v_ground * sr0


def crop(az_size, rg_span, az_span, rg_sampling, inc_angle, v_ground, prf, avg_az_shift,
         std_az_shift):
    """insar_processor.py:156-163: (az_min, az_max, rg_min, rg_max, az_guard,
    enough) with enough False where the reference prints 'Not enough
    edge-effect free image' and returns."""
    rg_min = 0
    rg_max = int(rg_span / (C / 2. / rg_sampling / np.sin(inc_angle)))
    az_min = int(az_size / 2. + (-az_span / 2. + avg_az_shift) / (v_ground / prf))
    az_max = int(az_size / 2. + (az_span / 2. + avg_az_shift) / (v_ground / prf))
    az_guard = int(std_az_shift / (v_ground / prf))
    return az_min, az_max, rg_min, rg_max, az_guard, not ((az_max - az_min) < (2 * az_guard - 10))


def shift_phasor(az_size, shift_dist):
    """insar_processor.py:169-171."""
    return np.exp(-2j * np.pi * shift_dist *
                  np.roll(np.arange(az_size) - az_size / 2, int(-az_size / 2)) / az_size)


def coregister(proc_data, b_ati, v_ground, prf):
    """insar_processor.py:165-174 on proc_data [ch][pol][az][rg] complex64
    (modified in place and returned)."""
    az_size = proc_data.shape[2]
    inter_chan_shift_dist = np.asarray(b_ati) / (v_ground / prf)
    for chind in range(proc_data.shape[0]):
        sa = shift_phasor(az_size, -inter_chan_shift_dist[chind]).reshape((1, az_size, 1))
        x = proc_data[chind].astype(np.complex128)
        proc_data[chind] = np.fft.ifft(np.fft.fft(x, axis=1) * sa, axis=1)
    return proc_data


def flattening_phasor(num_ch, rg_size, sr0, rg_sampling, alt, k0, b_xti, inc_angle, sr_to_geo):
    """insar_processor.py:177-184: [num_ch][1][1][rg] complex128. sr_to_geo
    is the geometry routine (utils/geometry.py:101-133), passed in."""
    slant_range = sr0 + np.arange(rg_size) / rg_sampling * C / 2
    _, inc, _, _ = sr_to_geo(slant_range, alt)
    flatearth_dinc = inc - inc_angle
    ph = np.zeros((num_ch, 1, 1, rg_size), dtype=complex)
    for ch in range(num_ch):
        ph[ch, 0, 0, :] = np.exp(1j * k0 * b_xti[ch] * flatearth_dinc)
    return ph


def l1b(proc_data, rg_ml, az_ml, ml_win, az_min, az_max, rg_min, rg_max):
    """insar_processor.py:186-217 on proc_data [ch][pol][az][rg]: returns
    (i_all [ch][pol][az'][rg'], cohs [pairs][az'][rg'] complex, coh_lut
    [C][C]) with C = ch * pol and the pairs in the reference's order."""
    ch_dim = proc_data.shape[0:2]
    p = proc_data.reshape([int(np.prod(ch_dim)), proc_data.shape[2], proc_data.shape[3]])
    i_all = []
    for c in range(p.shape[0]):
        this_i = smooth(smooth(np.abs(p[c]) ** 2., rg_ml, axis=1, window=ml_win),
                        az_ml, axis=0, window=ml_win)
        i_all.append(this_i[az_min:az_max, rg_min:rg_max])
    i_all = np.array(i_all)
    cohs = []
    tind = 0
    coh_lut = np.zeros((p.shape[0], p.shape[0]), dtype=int) - 1
    for c1 in range(p.shape[0]):
        for c2 in range(c1 + 1, p.shape[0]):
            coh_lut[c1, c2] = tind
            tind += 1
            t = smooth(smooth(p[c2] * np.conj(p[c1]), rg_ml, axis=1, window=ml_win),
                       az_ml, axis=0, window=ml_win)
            cohs.append(t[az_min:az_max, rg_min:rg_max] / np.sqrt(i_all[c1] * i_all[c2]))
    i_all = i_all.reshape(ch_dim + (az_max - az_min, rg_max - rg_min))
    return i_all, np.array(cohs), coh_lut


def ati_stats(i_all, cohs, coh_lut, b_ati, f0, v_ground, az_sampling, inc_angle, rg_sampling,
              rg_ml, az_ml, num_ch, vx, vz, dx, dy, sr0):
    """The numbers ati_process writes (ati_processor.py:97-470), from the L1b
    products and the float32 surface velocities at t = 0. inc_angle in
    radians; sr0 as the reference recomputes it (inc_to_sr, :143)."""
    k0 = 2. * np.pi * f0 / C
    grg = C / 2. / rg_sampling / np.sin(inc_angle)
    azg = v_ground / az_sampling
    res_fact = np.ceil(np.sqrt((grg / dx) * (azg / dy)))
    out = {'grg_grid_spacing': grg, 'az_grid_spacing': azg}
    v_r = radial_velocity(vx, vz, inc_angle)
    v_r_ml = smooth(smooth(v_r, res_fact * rg_ml, axis=1), res_fact * az_ml, axis=0)
    out['v_radial_surf_mean'] = np.mean(v_r)
    out['v_radial_surf_std'] = np.std(v_r)
    out['v_radial_surf_ml_std'] = np.std(v_r_ml)
    v_h_ml = smooth(smooth(vx, res_fact * rg_ml, axis=1), res_fact * az_ml, axis=0)
    out['v_horizo_surf_mean'] = np.mean(vx)
    out['v_horizo_surf_std'] = np.std(vx)
    out['v_horizo_surf_ml_std'] = np.std(v_h_ml)
    std_az_shift = out['v_radial_surf_std'] / v_ground * sr0
    az_guard = int(std_az_shift / (v_ground / az_sampling))
    out['az_guard'] = az_guard
    npol = i_all.shape[1]
    tau_ati = np.asarray(b_ati) / v_ground
    if num_ch > 1:
        phases = np.array([uwphase(cohs[coh_lut[0, p, 1, p]]) for p in range(npol)])
        v_rad = -phases / tau_ati[1] / (k0 * 2.)
        v_hor = -phases / tau_ati[1] / (k0 * 2.) / np.sin(inc_angle)
        v_rad = v_rad[:, az_guard:-az_guard, 5:-5]
        v_hor = v_hor[:, az_guard:-az_guard, 5:-5]
        out['v_radial_est_mean'] = [np.mean(v_rad[p]) for p in range(npol)]
        out['v_radial_est_std'] = [np.std(v_rad[p]) for p in range(npol)]
        sm = [smooth(smooth(v_rad[p], rg_ml, axis=1), az_ml, axis=0) for p in range(npol)]
        out['v_radial_est_ml_mean'] = [np.mean(s) for s in sm]
        out['v_radial_est_ml_std'] = [np.std(s) for s in sm]
        out['v_horizo_est_mean'] = [np.mean(v_hor[p]) for p in range(npol)]
        out['v_horizo_est_std'] = [np.std(v_hor[p]) for p in range(npol)]
        c = np.array(cohs)[:, az_guard:-az_guard, 5:-5]
        out['v_r_dop'] = np.mean(np.mean(v_rad, axis=-1), axis=-1)
        out['coh_mean'] = np.mean(np.mean(c, axis=-1), axis=-1)
        out['abscoh_mean'] = np.mean(np.mean(np.abs(c), axis=-1), axis=-1)
    out['nrcs'] = 10 * np.log10(np.mean(np.mean(i_all[:, :, az_guard:-az_guard, 5:-5], axis=-1),
                                        axis=-1))
    return out
