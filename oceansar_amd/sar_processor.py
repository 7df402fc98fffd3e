"""SAR focusing of the raw data file: sar_focus() on the GPU.

Mirrors oceansar/radarsim/sar_processor.py:30-260 (same signature, same
inputs and output file): the cfg's [processing] and [sar] keys, the raw file
read through RawFile.get as the reference reads it (:71-77), per channel the
range / azimuth processing loop (:131-225) in osar_sar_focus (four FP64 FFT
passes on the device, complex128 as the reference's numpy arrays), and the
SLC written with the reference's ProcFile layout (:237-254). The 1-D tables
of the loop (range and Doppler frequencies, antenna pattern, range migration,
azimuth weighting and matched filter) are computed on the host with the
reference's expressions. Plots are not produced (no plotting dependency).
"""
import ctypes

import numpy as np

from . import _lib
from . import constants as const
from . import geometry as geo
from . import raw_io
from .antenna import sinc_bp
from .cfg import ConfigFile
from .utils import optimize_fftsize


def sinc_1tx_nrx(sin_angle, L, f0, num_chan, field=False):
    """oceansar/radarsim/antenna.py:28-43: its `field` reaches sinc_bp as a
    positional *dummy, so the pattern is the power pattern whatever `field` is."""
    bp_tx = sinc_bp(sin_angle, L, f0)
    bp_rx = sinc_bp(sin_angle, L / num_chan, f0)
    return bp_tx * bp_rx


class FocusParams:
    """The scalars of the processing loop (sar_processor.py:44-67, 71-77, 155-163)."""

    def __init__(self, prf, rg_sampling, f0, v_ground, sr0, doppler_bw, az_weighting,
                 ant_l_tx=None, ant_l_rx=None, ant_l=None, l_total=False, num_ch=1):
        self.prf, self.rg_sampling, self.f0 = float(prf), float(rg_sampling), float(f0)
        self.v_ground, self.sr0 = float(v_ground), float(sr0)
        self.doppler_bw, self.az_weighting = float(doppler_bw), float(az_weighting)
        self.ant_l_tx, self.ant_l_rx, self.ant_l = ant_l_tx, ant_l_rx, ant_l
        self.l_total, self.num_ch = bool(l_total), int(num_ch)

    @classmethod
    def from_cfg(cls, cfg, sr0):
        sar, proc = cfg.sar, cfg.processing
        v_ground = sar.v_ground
        if v_ground == 'auto':
            v_ground = geo.orbit_to_vel(sar.alt, ground=True)
        kw = {}
        if hasattr(sar, 'ant_L'):
            kw.update(ant_l=sar.ant_L, l_total=getattr(sar, 'L_total', False))
        else:
            kw.update(ant_l_tx=sar.ant_L_tx, ant_l_rx=sar.ant_L_rx)
        return cls(sar.prf, sar.rg_bw * sar.over_fs, sar.f0, v_ground, sr0, proc.doppler_bw,
                   proc.az_weighting, num_ch=sar.num_ch, **kw)

    def tables(self, n_az, n_rg):
        """fr [n_rg], rcmc_fa [n_az] and the azimuth filter h [n_az] (complex128):
        sar_processor.py:148-213 with the padded sizes."""
        l0 = const.c / self.f0
        fr = np.fft.fftfreq(n_rg, 1 / self.rg_sampling)
        fa = np.fft.fftfreq(n_az, 1 / self.prf)
        sin_az = fa * l0 / (2 * self.v_ground)
        if self.ant_l is not None:
            if self.l_total:
                beam = sinc_1tx_nrx(sin_az, self.ant_l * self.num_ch, self.f0, self.num_ch)
            else:
                beam = sinc_1tx_nrx(sin_az, self.ant_l, self.f0, 1)
        else:
            beam = (sinc_bp(sin_az, self.ant_l_tx, self.f0, field=True) *
                    sinc_bp(sin_az, self.ant_l_rx, self.f0, field=True))
        rcmc_fa = self.sr0 / np.sqrt(1 - (fa * (l0 / 2.) / self.v_ground) ** 2.) - self.sr0
        n_samp = 2 * (int(self.doppler_bw / (fa[1] - fa[0])) / 2)
        w = self.az_weighting
        weighting = w - (1. - w) * np.cos(2 * np.pi * np.linspace(0, 1., int(n_samp)))
        L_win = np.sum(np.abs(weighting) ** 2) / weighting.size
        weighting /= np.sqrt(L_win)
        if fa.size > n_samp:
            zeros = np.zeros(n_az)
            zeros[0:int(n_samp)] = weighting
            weighting = np.roll(zeros, int(-n_samp / 2))
        elif weighting.size != n_az:
            raise ValueError('doppler_bw %g Hz exceeds the PRF band of %d samples'
                             % (self.doppler_bw, n_az))
        with np.errstate(divide='ignore', invalid='ignore'):
            weighting = np.where(np.abs(beam) > 0, weighting / beam, 0)
        ph_ac = 4. * np.pi / l0 * self.sr0 * (np.sqrt(1. - (fa * l0 / 2. / self.v_ground) ** 2.)
                                              - 1.)
        h = np.exp(1j * ph_ac) * weighting
        return fr, rcmc_fa, h

    def valid_rows(self, az_size_orig):
        """[lo, hi) of the lines kept after focusing (sar_processor.py:222-225)."""
        l0 = const.c / self.f0
        n_val_az_2 = np.floor(self.doppler_bw / 2. / (2. * self.v_ground ** 2. / l0 / self.sr0)
                              * self.prf / 2.) * 2.
        return int(n_val_az_2), int(az_size_orig - n_val_az_2 - 1)


def focus_channel(ctx, raw_ch, params, work=None):
    """SLC [pol][kept lines][rg] (complex128 device tensor) of one channel.

    raw_ch: complex64 device tensor [pol][az][rg] (a channel slice of the raw
    cube [pol][ch][az][rg] works as is: its planes are raw_pol_stride apart).
    work: optional complex128 device workspace of at least pol * n_az * n_rg.
    """
    import torch
    if raw_ch.dtype != torch.complex64 or raw_ch.dim() != 3:
        raise TypeError('raw_ch must be a complex64 [pol][az][rg] device tensor')
    npol, az0, rg0 = raw_ch.shape
    if raw_ch.stride(2) != 1 or raw_ch.stride(1) != rg0:
        raise TypeError('raw_ch lines must be contiguous')
    n_az, n_rg = (int(v) for v in optimize_fftsize((az0, rg0)))
    fr, rcmc, h = params.tables(n_az, n_rg)
    lo, hi = params.valid_rows(az0)
    if not (0 <= lo < hi <= az0):
        raise ValueError('no valid lines: %d raw lines, %d removed at each end' % (az0, lo))
    dev = raw_ch.device
    fr_d = torch.as_tensor(fr, dtype=torch.float64, device=dev)
    rcmc_d = torch.as_tensor(rcmc, dtype=torch.float64, device=dev)
    h_d = torch.as_tensor(h, dtype=torch.complex128, device=dev)
    if work is None or work.numel() < npol * n_az * n_rg:
        work = torch.empty(npol * n_az * n_rg, dtype=torch.complex128, device=dev)
    slc = torch.empty((npol, hi - lo, rg0), dtype=torch.complex128, device=dev)
    d = _lib.FocusDesc()
    d.n_pol, d.n_az0, d.n_rg0, d.n_az, d.n_rg = npol, az0, rg0, n_az, n_rg
    d.az_lo, d.az_hi = lo, hi
    d.raw_pol_stride = raw_ch.stride(0)
    d.raw = raw_ch.data_ptr()
    d.fr, d.rcmc, d.h = fr_d.data_ptr(), rcmc_d.data_ptr(), h_d.data_ptr()
    d.work, d.slc = work.data_ptr(), slc.data_ptr()
    _lib.check(ctx.lib.osar_sar_focus(ctx.h, ctypes.byref(d), _lib.stream_ptr()))
    # the tables must live until the stream has run the passes
    torch.cuda.current_stream(dev).synchronize()
    return slc


def sar_focus(cfg_file, raw_output_file, output_file, device=0):
    """Focus the raw file and write the SLC file (sar_processor.py:30-260)."""
    import torch
    cfg = ConfigFile(cfg_file)
    with raw_io.RawFile(raw_output_file, 'r') as raw_file:
        raw_data = raw_file.get('raw_data*')
        sr0 = raw_file.get('sr0')
        inc_angle = raw_file.get('inc_angle')
        b_ati = raw_file.get('b_ati')
        b_xti = raw_file.get('b_xti')
    params = FocusParams.from_cfg(cfg, sr0)
    num_ch = int(cfg.sar.num_ch)
    ctx = _lib.context(device)
    raw_d = torch.as_tensor(np.ascontiguousarray(raw_data), device=torch.device('cuda', device))
    slc = []
    work = None
    for ch in range(num_ch):
        s = focus_channel(ctx, raw_d[:, ch], params, work)
        slc.append(s.cpu().numpy())
    slc = np.array(slc, dtype=complex)
    with raw_io.ProcFile(output_file, 'w', slc.shape) as pf:
        pf.set('slc*', slc)
        pf.set('inc_angle', inc_angle)
        pf.set('f0', params.f0)
        pf.set('num_ch', num_ch)
        pf.set('ant_L', params.ant_l_tx if params.ant_l is None else params.ant_l)
        pf.set('prf', params.prf)
        pf.set('v_ground', params.v_ground)
        pf.set('orbit_alt', cfg.sar.alt)
        pf.set('sr0', sr0)
        pf.set('rg_sampling', cfg.sar.rg_bw * cfg.sar.over_fs)
        pf.set('rg_bw', cfg.sar.rg_bw)
        pf.set('b_ati', b_ati)
        pf.set('b_xti', b_xti)
    return slc
