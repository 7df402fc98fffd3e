"""InSAR L1b processing of the focused SLC file on the GPU: insar_process().

Mirrors oceansar/radarsim/insar_processor.py:42-237 (same signature, cfg keys,
input files and L1b output layout; no plots): the SLC read through
ProcFile.get as complex64 (:80-94), the surface velocities at t = 0 that set
the expected azimuth shift and the region of interest (:97-163), then on the
device the per-channel sub-sample azimuth shift and flat-earth removal
(osar_insar_coregister, :165-185) and the multilooked intensities,
interferograms and coherences (osar_insar_multilook, :186-214), written with
the reference's L1bFile layout (:215-236).

Decisions (the reference's arithmetic depends on the numpy of its era,
numpy < 1.24, whose np.complex it uses):
  - the float32 surface velocities times the float64 sin/cos of the incidence
    stay float32 (value-based casting), and their mean and standard deviation
    are float32 reductions, as there;
  - the FFTs of the complex64 SLC are complex128 (numpy < 2 computes every FFT
    in double), and the shifted values are rounded to complex64 where the
    reference assigns them back into its complex64 array;
  - with flatten = False the reference keeps complex64 images and a flat
    window sums |s|^2 and the interferograms in single precision; here every
    sum is FP64 (with flatten = True, the default, both are FP64 and flat
    windows give the reference's sums bit for bit);
  - the smoothed surface velocities (:127, :134) feed only the reference's
    plots and are not computed.
"""
import ctypes

import numpy as np

from . import _lib
from . import constants as const
from . import raw_io
from .cfg import ConfigFile
from .geometry import sr_to_geo
from .ocean import OceanSurface

WINDOWS = ('flat', 'hanning', 'hamming', 'bartlett', 'blackman')


def window_spec(window, window_len):
    """(length, flat, taps) of one axis of utils.smooth (misc.py:148-197) as
    the processors call it (axis given): flat windows are smooth1d's moving
    average; other windows shorter than 3 leave the data as they are (one
    unit tap); longer ones are numpy's window normalised by its sum."""
    if window not in WINDOWS:
        raise ValueError('Window type not supported')
    L = int(window_len)
    if window == 'flat':
        if L < 1:
            raise ValueError('flat window of length %d' % L)
        return L, 1, None
    if L < 3:
        return 1, 0, np.ones(1)
    w = getattr(np, window)(L)
    return L, 0, w / w.sum()


def shift_kernel(az_size, shift_dist):
    """(n_fft, spectrum) for osar_insar_coregister: the reference's shift
    phasor S (insar_processor.py:169-171) as the circular kernel h = IFFT(S),
    wrapped into a power-of-two length M >= 2 az_size - 1 (g[j] = h[j],
    g[M - j] = h[az_size - j]) and transformed, FFT(g) / M."""
    S = np.exp(-2j * np.pi * shift_dist *
               np.roll(np.arange(az_size) - az_size / 2, int(-az_size / 2)) / az_size)
    h = np.fft.ifft(S)
    M = 16
    while M < 2 * az_size - 1:
        M *= 2
    g = np.zeros(M, dtype=np.complex128)
    g[:az_size] = h
    if az_size > 1:
        g[M - az_size + 1:] = h[1:]
    return M, np.fft.fft(g) / M


def flat_earth_phasor(rg_size, sr0, rg_sampling, alt, k0, b_xti, inc_angle):
    """exp(i k0 b_xti (inc(r) - inc0)) per range sample (insar_processor.py:177-184)."""
    slant_range = sr0 + np.arange(rg_size) / rg_sampling * const.c / 2
    _, inc, _, _ = sr_to_geo(slant_range, alt)
    return np.exp(1j * k0 * b_xti * (inc - inc_angle))


class CoregPlan:
    """Device tables of one channel's shift: the kernel spectrum for n_az
    lines and shift_dist lines, and the flat-earth phasor (None: none)."""

    def __init__(self, n_az, shift_dist, flat=None, device=0):
        import torch
        dev = torch.device('cuda', device) if isinstance(device, int) else device
        self.n_az = int(n_az)
        self.n_fft, G = shift_kernel(self.n_az, shift_dist)
        self.kernel = torch.as_tensor(G, dtype=torch.complex128, device=dev)
        self.flat = None if flat is None else torch.as_tensor(
            np.asarray(flat, dtype=np.complex128), device=dev)


def coregister_channel(ctx, slc, plan, out=None):
    """Shift one channel's planes slc (complex64 device tensor [pol][az][rg])
    along azimuth and flatten them as the CoregPlan says, on the current
    stream. Returns the complex128 device tensor [pol][az][rg] (out, when
    given)."""
    import torch
    if slc.dtype != torch.complex64 or slc.dim() != 3:
        raise TypeError('slc must be a complex64 [pol][az][rg] device tensor')
    npol, naz, nrg = slc.shape
    if slc.stride(2) != 1 or slc.stride(1) != nrg:
        raise TypeError('slc lines must be contiguous')
    if naz != plan.n_az:
        raise ValueError('plan for %d lines, slc has %d' % (plan.n_az, naz))
    if plan.flat is not None and plan.flat.numel() != nrg:
        raise ValueError('flat phasor of %d samples for %d range samples'
                         % (plan.flat.numel(), nrg))
    if out is None:
        out = torch.empty((npol, naz, nrg), dtype=torch.complex128, device=slc.device)
    elif out.dtype != torch.complex128 or tuple(out.shape) != (npol, naz, nrg) \
            or not out.is_contiguous():
        raise TypeError('out must be a contiguous complex128 [pol][az][rg] tensor')
    d = _lib.CoregDesc()
    d.n_pol, d.n_az, d.n_rg, d.n_fft = npol, naz, nrg, plan.n_fft
    d.in_pol_stride = slc.stride(0)
    d.in_, d.kernel = slc.data_ptr(), plan.kernel.data_ptr()
    d.flat = None if plan.flat is None else plan.flat.data_ptr()
    d.out = out.data_ptr()
    _lib.check(ctx.lib.osar_insar_coregister(ctx.h, ctypes.byref(d), _lib.stream_ptr()))
    return out


def multilook(ctx, images, rg_ml, az_ml, window='flat', crop=None):
    """Smooth images [C][az][rg] (complex128: intensities, interferograms of every
    pair c1 < c2 and their coherence; float64: the images themselves) along
    range then azimuth with utils.smooth's window of each length, on rows
    [az_lo, az_hi) and columns [rg_lo, rg_hi) = crop (default: all). Returns
    (intensity [C][rows][cols], coherence [pairs][..], phase [pairs][..]) as
    float64 device tensors (None for real images)."""
    import torch
    if images.dim() != 3 or not images.is_contiguous():
        raise TypeError('images must be a contiguous [C][az][rg] device tensor')
    if images.dtype == torch.complex128:
        is_complex = 1
    elif images.dtype == torch.float64:
        is_complex = 0
    else:
        raise TypeError('images must be complex128 or float64')
    C, naz, nrg = images.shape
    az_lo, az_hi, rg_lo, rg_hi = (0, naz, 0, nrg) if crop is None else crop
    Lr, fr, tr = window_spec(window, rg_ml)
    La, fa, ta = window_spec(window, az_ml)
    dev = images.device
    d = _lib.MultilookDesc()
    d.n_img, d.n_az, d.n_rg, d.is_complex = C, naz, nrg, is_complex
    d.rg_len, d.az_len, d.rg_flat, d.az_flat = Lr, La, fr, fa
    d.az_lo, d.az_hi, d.rg_lo, d.rg_hi = az_lo, az_hi, rg_lo, rg_hi
    d.img_stride = naz * nrg
    keep = []
    for name, taps in (('rg_taps', tr), ('az_taps', ta)):
        if taps is not None:
            t = torch.as_tensor(taps, dtype=torch.float64, device=dev)
            keep.append(t)
            setattr(d, name, t.data_ptr())
    nwork = ctx.lib.osar_multilook_work_bytes(ctypes.byref(d))
    if nwork < 0:
        raise ValueError('multilook: empty crop [%d, %d) x [%d, %d)' % crop)
    work = torch.empty(max(1, nwork // 8), dtype=torch.float64, device=dev)
    rows, cols = az_hi - az_lo, rg_hi - rg_lo
    pairs = C * (C - 1) // 2 if is_complex else 0
    inten = torch.empty((C, rows, cols), dtype=torch.float64, device=dev)
    coh = torch.empty((pairs, rows, cols), dtype=torch.float64, device=dev) if pairs else None
    pha = torch.empty((pairs, rows, cols), dtype=torch.float64, device=dev) if pairs else None
    d.img, d.work, d.intensity = images.data_ptr(), work.data_ptr(), inten.data_ptr()
    if pairs:
        d.coherence, d.phase = coh.data_ptr(), pha.data_ptr()
    # taps and workspace are freed to torch's stream-ordered cache: later work
    # that reuses them is queued behind these launches
    _lib.check(ctx.lib.osar_insar_multilook(ctx.h, ctypes.byref(d), _lib.stream_ptr()))
    del keep, work
    return inten, coh, pha


def slice_range(lo, hi, n, what):
    """[lo, hi) as numpy's a[lo:hi] selects it on an axis of n; the reference
    then reshapes to hi - lo samples, which fails unless the slice has them."""
    a, b, _ = slice(lo, hi).indices(n)
    if b - a != hi - lo or hi <= lo:
        raise ValueError('%s [%d, %d) does not lie in [0, %d): the reference cannot reshape '
                         'its products to %d samples' % (what, lo, hi, n, hi - lo))
    return a, b


def surface_velocity_stats(surface, inc_angle, v_ground, sr0):
    """Radial surface velocity statistics (insar_processor.py:126-142) from the
    float32 fields of a surface at t = 0: (mean, std, avg_az_shift,
    std_az_shift); float32 arithmetic as the reference's numpy."""
    v_r = (surface.Vx * np.float32(np.sin(inc_angle)) -
           surface.Vz * np.float32(np.cos(inc_angle)))
    mean, std = np.mean(v_r), np.std(v_r)
    return mean, std, -mean / v_ground * sr0, std / v_ground * sr0


def insar_process(cfg_file, proc_output_file, ocean_file, output_file, device=0):
    """insar_processor.py:42-237. Returns (i_all, cohs, coh_lut) as host arrays,
    or None where the reference returns early (not enough edge-effect free
    image)."""
    import torch
    cfg = ConfigFile(cfg_file)
    rg_ml, az_ml = cfg.insar.rg_ml, cfg.insar.az_ml
    ml_win = cfg.insar.ml_win
    flatten = cfg.insar.flatten
    with raw_io.ProcFile(proc_output_file, 'r') as pc:
        proc_data = pc.get('slc*')
        sr0 = pc.get('sr0')
        b_ati = np.atleast_1d(pc.get('b_ati'))
        b_xti = np.atleast_1d(pc.get('b_xti'))
        f0 = pc.get('f0')
        prf = pc.get('prf')
        num_ch = int(pc.get('num_ch'))
        rg_sampling = pc.get('rg_sampling')
        v_ground = pc.get('v_ground')
        alt = pc.get('orbit_alt')
        inc_angle = np.deg2rad(pc.get('inc_angle'))
    k0 = 2. * np.pi * f0 / const.c

    surface = OceanSurface(device)
    surface.load(ocean_file, compute=['D', 'V'])
    surface.t = 0.
    _, _, avg_az_shift, std_az_shift = surface_velocity_stats(surface, inc_angle, v_ground, sr0)

    nch, npol, az_size, rg_size = proc_data.shape
    rg_min = 0
    rg_max = int(surface.Lx / (const.c / 2. / rg_sampling / np.sin(inc_angle)))
    az_min = int(az_size / 2. + (-surface.Ly / 2. + avg_az_shift) / (v_ground / prf))
    az_max = int(az_size / 2. + (surface.Ly / 2. + avg_az_shift) / (v_ground / prf))
    az_guard = int(std_az_shift / (v_ground / prf))
    if (az_max - az_min) < (2 * az_guard - 10):
        print('Not enough edge-effect free image')
        return None
    a0, a1 = slice_range(az_min, az_max, az_size, 'azimuth crop')
    r0, r1 = slice_range(rg_min, rg_max, rg_size, 'range crop')

    ctx = _lib.context(device)
    dev = torch.device('cuda', device)
    slc = torch.as_tensor(np.ascontiguousarray(proc_data), device=dev)
    images = torch.empty((nch * npol, az_size, rg_size), dtype=torch.complex128, device=dev)
    inter_chan_shift_dist = b_ati / (v_ground / prf)
    for ch in range(nch):
        flat = (flat_earth_phasor(rg_size, sr0, rg_sampling, alt, k0, b_xti[ch], inc_angle)
                if flatten else None)
        plan = CoregPlan(az_size, -inter_chan_shift_dist[ch], flat, dev)
        coregister_channel(ctx, slc[ch], plan, out=images[ch * npol:(ch + 1) * npol])
    del slc
    inten, coh, pha = multilook(ctx, images, rg_ml, az_ml, ml_win, (a0, a1, r0, r1))
    del images
    C = nch * npol
    i_all = inten.cpu().numpy().reshape((nch, npol, a1 - a0, r1 - r0))
    coh_lut = np.zeros((C, C), dtype=int) - 1
    tind = 0
    for c1 in range(C):
        for c2 in range(c1 + 1, C):
            coh_lut[c1, c2] = tind
            tind += 1
    ml_coh = coh.cpu().numpy() if coh is not None else np.zeros((0, a1 - a0, r1 - r0))
    ml_pha = pha.cpu().numpy() if pha is not None else np.zeros((0, a1 - a0, r1 - r0))
    with raw_io.L1bFile(output_file, 'w', i_all.shape) as lf:
        lf.set('ml_intensity', i_all)
        if C > 1:
            lf.set('ml_coherence', ml_coh)
            lf.set('ml_phase', ml_pha)
        lf.set('coh_lut', coh_lut.reshape((num_ch, npol, num_ch, npol)))
        lf.set('inc_angle', inc_angle)
        lf.set('f0', f0)
        lf.set('num_ch', num_ch)
        lf.set('az_sampling', prf)
        lf.set('v_ground', v_ground)
        lf.set('orbit_alt', alt)
        lf.set('sr0', sr0)
        lf.set('rg_sampling', rg_sampling)
        lf.set('b_ati', b_ati)
        lf.set('b_xti', b_xti)
        lf.set('rg_ml', rg_ml)
        lf.set('az_ml', az_ml)
    return i_all, ml_coh * np.exp(1j * ml_pha), coh_lut
