"""Device-resident scene and the GPU form of the sar_raw pulse loop.

DeviceScene uploads the host set-up of a SarScene once (geometry tables,
closure series, spectrum table, range-response table, wave coefficients) and
drives the three kernels of liboceansar_hip.so: surface synthesis
(osar_surface_synth), Bragg facets (osar_facet_bragg) and echo accumulation
(osar_echo_accumulate). torch is used only to own device buffers.
"""
import ctypes
import math

import numpy as np

from . import _lib

POL_CODE = {'vv': 0, 'hh': 1, 'DP': 2}


def _t(a, dtype, device):
    import torch
    if torch.is_tensor(a):   # already on the device (osar_surface_init coefficients)
        return a.to(device=device, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)


def mirror_grid(k):
    """True when the float32 wave numbers k form a 2 pi fftfreq grid as the
    synthesis row pass uses it: k[0] = 0 and k[n - i] = -k[i] for i != n / 2."""
    k = np.asarray(k, dtype=np.float32)
    n = len(k)
    i = np.arange(n)
    ok = k[(n - i) % n] == -k[i]
    ok[n // 2] = True
    return bool(ok.all())


def synth_fields(ctx, surf, kx, ky, coefs, times, out, fields=None, mirror=None,
                 slope_fp32=False):
    """osar_surface_synth: fields of surface `surf` (a SurfaceInit) at `times`.

    kx, ky, coefs: device tensors (float32 [nx], [ny], complex64 [ny][nx]).
    out[name]: float32 device tensor [>= n_times][nx][ny] (x-major) per field.
    `times` is a sequence of floats (copied to the device) or a contiguous float64
    device tensor (used in place, no host round trip). Returns the device tensor
    of times, which must stay alive until the stream has run the launch.
    slope_fp32: the Diffx/Diffy spectrum round-trips through a complex64
    workspace (osar_surface_desc.slope_fp32).
    """
    import torch
    fields = fields or list(out.keys())
    if getattr(surf, 'swell_enable', False):
        raise NotImplementedError('monochromatic swell is not supported on the GPU path')
    if torch.is_tensor(times):
        if times.dtype != torch.float64 or not times.is_contiguous() or times.device != coefs.device:
            raise TypeError('times must be a contiguous float64 tensor on the surface device')
        times_t = times
    else:
        times_t = torch.as_tensor(np.asarray(times, dtype=np.float64), device=coefs.device)
    d = _lib.SurfaceDesc()
    d.nx, d.ny = surf.Nx, surf.Ny
    mask = 0
    for n in fields:
        mask |= _lib.FIELD_BITS[n]
    d.fields = mask
    d.choppy = 1 if surf.choppy else 0
    d.n_times = len(times)
    if mirror is None:   # checked on the host copies of the uploaded tables
        mirror = mirror_grid(surf.kx_s) and mirror_grid(surf.ky_s)
    d.mirror_grid = 1 if mirror else 0
    d.slope_fp32 = 1 if slope_fp32 else 0
    d.depth = -1.0 if surf.depth is None else float(surf.depth)
    d.current_x, d.current_y = float(surf.current[0]), float(surf.current[1])
    d.kx_res, d.ky_res = float(surf.kx_res), float(surf.ky_res)
    d.kx, d.ky, d.coefs = _lib.ptr(kx), _lib.ptr(ky), _lib.ptr(coefs)
    d.times = _lib.ptr(times_t)
    for i, n in enumerate(_lib.FIELD_ORDER):
        if n in fields:
            t = out[n]
            if t.dtype != torch.float32 or not t.is_contiguous():
                raise TypeError('field %s must be a contiguous float32 tensor' % n)
            if tuple(t.shape[-2:]) != (surf.Nx, surf.Ny) or t.shape[0] < len(times):
                raise TypeError('field %s has shape %s, need [>=%d, %d, %d]'
                                % (n, tuple(t.shape), len(times), surf.Nx, surf.Ny))
            d.out[i] = t.data_ptr()
    _lib.check(ctx.lib.osar_surface_synth(ctx.h, ctypes.byref(d), _lib.stream_ptr()))
    return times_t


def scene_fields(sc):
    """The surface fields the pulse loop of scene `sc` reads, in field order."""
    f = ['Dx', 'Dy', 'Dz']
    if sc.bragg_enable:
        f += ['Diffx', 'Diffy']
        if getattr(sc, 'use_hmtf', False):
            f += ['hMTF']
    if sc.spec_enable:
        f += ['Diffx', 'Diffy', 'Diffxx', 'Diffyy', 'Diffxy']
    return [n for n in _lib.FIELD_ORDER if n in f]


def spectra_bytes(fields, choppy, slope_fp32):
    """Workspace bytes per grid point and snapshot of osar_surface_synth for
    these fields: 16 per complex128 spectrum, 8 per complex64 one (the plan of
    osar_surface_synth: monomial fields in pairs, choppy slopes alone, five
    spectra for choppy curvatures, M_h taking an odd field out)."""
    fields = set(fields)
    curv = bool(fields & {'Diffxx', 'Diffyy', 'Diffxy'})
    choppy2 = choppy and curv
    n128 = 5 if choppy2 else 0
    items = [n for n in _lib.FIELD_ORDER[:14] if n in fields]
    if choppy2:
        items = [n for n in items if n not in ('Diffx', 'Diffy', 'Diffxx', 'Diffyy', 'Diffxy')]
    elif choppy:
        n128 += sum(1 for n in ('Diffx', 'Diffy') if n in items)
        items = [n for n in items if n not in ('Diffx', 'Diffy')]
    n64 = 0
    if slope_fp32 and not choppy:
        lo = [n for n in items if n in ('Diffx', 'Diffy')]
        items = [n for n in items if n not in ('Diffx', 'Diffy')]
        n64 = 1 if lo else 0
    m = len(items)
    if 'hMTF' in fields:
        n128 += 1
        m -= m % 2
    n128 += (m + 1) // 2
    return 16 * n128 + 8 * n64


def arange_grid(k):
    """(k0, dk, 1) when k[i] == k0 + i * dk exactly for every i, with the product
    rounded before the sum (numpy.arange fills its output that way), so the Bragg
    kernel may compute grid points instead of loading them; (k0, 0, 0) otherwise."""
    k = np.asarray(k, dtype=np.float64)
    if k.size < 2:
        return (float(k[0]) if k.size else 0.), 0., 0
    dk = k[1] - k[0]
    exact = k[0] + np.arange(k.size, dtype=np.float64) * dk
    if dk > 0 and np.array_equal(exact, k):
        return float(k[0]), float(dk), 1
    return float(k[0]), 0., 0


class RunState:
    """Buffers and position of one DeviceScene run (DeviceScene.start/step/finish)."""


class LaunchShape:
    """Echo-accumulator launch shape: pulses per workgroup, rows per workgroup, LDS
    window, columns per workgroup (0 = all)."""

    def __init__(self, pulses_per_block=0, rows_per_block=0, window_bins=0, cols_per_block=0):
        self.pulses_per_block = pulses_per_block
        self.rows_per_block = rows_per_block
        self.window_bins = window_bins
        self.cols_per_block = cols_per_block

    def as_list(self):
        return [self.pulses_per_block, self.rows_per_block, self.window_bins,
                self.cols_per_block]

    def __repr__(self):
        s = 'LaunchShape(P=%d, R=%d, W=%d' % (self.pulses_per_block, self.rows_per_block,
                                            self.window_bins)
        if self.cols_per_block:
            s += ', C=%d' % self.cols_per_block
        return s + ')'


class DeviceScene:
    def __init__(self, sc, device=0, lds_budget=160 * 1024, ctx=None, shape=None):
        """Upload the scene's tables. `ctx`: the library context to launch with
        (each concurrently used stream needs its own: a context owns the
        synthesis workspace); default the device's shared context. `shape`: a
        LaunchShape for the accumulator instead of default_shape()."""
        import torch
        self.sc = sc
        self.ctx = ctx if ctx is not None else _lib.context(device)
        self.dev = torch.device('cuda', device)
        dev = self.dev
        f64, f32, i32, c64 = torch.float64, torch.float32, torch.int32, torch.complex64
        self.sr_col = _t(sc.sr, f64, dev)
        self.sin_inc = _t(sc.sin_inc, f64, dev)
        self.cos_inc = _t(sc.cos_inc, f64, dev)
        self.inc_col = _t(sc.inc, f64, dev)
        self.dinc_col = _t(sc.inc - sc.inc_angle, f64, dev)
        self.y_row = _t(sc.y64, f64, dev)
        self.sinc = _t(sc.sinc_vec, f32, dev)
        self.npol = len(sc.pols)
        self.n_pulses = sc.az_steps
        if sc.bragg_enable:
            self.cap_rate = _t(sc.cap_rate, f64, dev)
            self.scat_m = _t(self._extend(sc.scat_m), c64, dev)
            self.scat_p = _t(self._extend(sc.scat_p), c64, dev)
            # x-major [nx][ny], like every per-facet device array
            # uploaded row-major, transposed on the device (a host transpose of a
            # 2048^2 table costs ~0.1 s)
            self.ind0_m = _t(np.mod(sc.scat_m.ind0, sc.scat_m.N).astype(np.int32), i32,
                             dev).t().contiguous()
            self.ind0_p = _t(np.mod(sc.scat_p.ind0, sc.scat_p.N).astype(np.int32), i32,
                             dev).t().contiguous()
            self.rcs_slot = _t(sc.rcs_slot, i32, dev)
            if np.any(sc.rcs_refresh % sc.K):
                raise ValueError('Bragg refresh pulses must be snapshot pulses (multiples of K)')
            self.refresh_pulse = _t(sc.rcs_refresh, i32, dev)
            self.refresh_snap = _t(sc.rcs_refresh // sc.K, i32, dev)
            self.spec_k = _t(sc.bragg.spec_k, f64, dev)
            self.spec_grid = arange_grid(sc.bragg.spec_k)
            self.spec_S = _t(sc.bragg.spec_S, f64, dev)
            # per-interval slopes of the linear interpolation, the same FP64
            # quotient the kernel would take per facet
            k, S = np.asarray(sc.bragg.spec_k, np.float64), np.asarray(sc.bragg.spec_S, np.float64)
            slope = np.zeros_like(S)
            slope[1:] = (S[1:] - S[:-1]) / (k[1:] - k[:-1])
            self.spec_slope = _t(slope, f64, dev)
        surf = sc.surface
        self.coefs = _t(surf.wave_coefs, c64, dev)
        self.kx = _t(surf.kx_s, f32, dev)
        self.ky = _t(surf.ky_s, f32, dev)
        n_snap = sc.az_steps // sc.K + 2
        self.snap_times = _t([sc.snapshot_time(m) for m in range(n_snap)], f64, dev)
        self.resid_max = self._resid_bound()
        self.sr_monotonic = bool(np.all(np.diff(sc.sr) >= 0))
        self.shape = shape if shape is not None else self.default_shape(lds_budget)
        # Bragg-only scenes read the slopes only through the Bragg amplitudes: their
        # spectrum takes the complex64 workspace (the specular stationary-phase test
        # keeps every field in complex128)
        self.slope_fp32 = bool(sc.bragg_enable and not sc.spec_enable)

    def fp64_tables(self):
        """complex128 scatterer series of the all-FP64 reference path (made once)."""
        import torch
        if not hasattr(self, 'scat64_m'):
            n = lambda rs: rs.s[np.arange(rs.N + self.n_pulses) % rs.N].astype(np.complex128)
            self.scat64_m = _t(n(self.sc.scat_m), torch.complex128, self.dev)
            self.scat64_p = _t(n(self.sc.scat_p), torch.complex128, self.dev)
        return self.scat64_m, self.scat64_p

    # ------------------------------------------------------------------
    def _extend(self, rs):
        """s_ext[i] = s[i mod N] for i < N + n_pulses, so the kernel indexes ind0 + p."""
        n = rs.N + self.n_pulses
        return rs.s[np.arange(n) % rs.N].astype(np.complex64)

    def _resid_bound(self):
        """Window margin for |residual range|: 8 x the spectral RMS elevation (|Dz| + |Dx|
        rarely exceed ~5 RMS together), plus the current drift.

        Only sizes the LDS window; facets beyond it take the global-memory path.
        """
        import torch
        n2 = float(self.sc.Nx * self.sc.Ny)
        c = self.coefs        # device complex64 coefficients
        power = float(torch.sum(c.real.double() ** 2 + c.imag.double() ** 2).item())
        rms = math.sqrt(power / 2.) / n2
        cur = float(np.hypot(*self.sc.surface.current)) * self.n_pulses * self.sc.t_step
        return 8. * rms + cur + 0.5

    def window_bins(self, rows, cols=0):
        """LDS window covering a tile of `rows` rows (and `cols` columns, 0 = all)
        at any pulse (see echo.hip)."""
        sc = self.sc
        az_max = max(abs(sc.az_now(0)), abs(sc.az_now(sc.az_steps))) + sc.surface.Ly
        mig_spread = rows * float(sc.surface.dy) * az_max / sc.sr0
        if cols and cols < sc.Nx:
            if not self.sr_monotonic:
                raise ValueError('column chunks need a non-decreasing slant-range table')
            lo = sc.sr[0:sc.Nx:cols]
            hi = sc.sr[np.minimum(np.arange(0, sc.Nx, cols) + cols, sc.Nx) - 1]
            sr_span = float(np.max(hi - lo))
        else:
            sr_span = float(sc.sr.max() - sc.sr.min())
        span = (sr_span + mig_spread + 2 * self.resid_max) / sc.binsize
        W = int(math.ceil(span)) + 8 + 6
        return min(((W + 15) // 16) * 16, ((sc.rg_samp + 15) // 16) * 16)

    @staticmethod
    def lds_bytes(P, n_ch, npol, W):
        """LDS of one accumulator workgroup (echo.hip: echo_lds_bytes, line_stride)."""
        nch_c = 1 if n_ch <= 1 else 2 if n_ch <= 2 else 4 if n_ch <= 4 else 8
        LS = ((2 * W + 31) // 32) * 32 + 4            # FP64 lines
        return 8 * npol * nch_c * P * LS + 4 * 168 + 4 * P   # lines, tap table, window starts

    def default_shape(self, lds_budget=160 * 1024, n_ch=None, pulses=(8, 4)):
        """The plain kernels are built for two workgroups per CU (4 waves per SIMD):
        the largest P in (8, 4) and the widest column chunk whose LDS lines fit
        twice in a CU (P = 8 loads each facet once for 8 pulses); otherwise
        (specular, XTI, or no fit) the largest P whose lines fit once. n_ch:
        size the lines for this many channels (a batch kernel built for the
        batch's largest channel count) instead of the scene's; pulses: the P
        values tried for two workgroups per CU."""
        sc = self.sc
        nch = sc.num_ch if n_ch is None else int(n_ch)
        if not (sc.spec_enable or np.any(sc.b_xti != 0)):
            for P in pulses:
                # one row group per wave (64 rows at P = 8): 256-row tiles measured
                # 0.8 % faster on the 2048^2 workload but read the surface fields
                # 1.6x over from HBM (a larger working set per XCD's L2)
                R = min(8 * (64 // P), sc.Ny)
                for cols in (sc.Nx, sc.Nx // 2, sc.Nx // 4, sc.Nx // 8):
                    if cols < 64 or (cols < sc.Nx and not self.sr_monotonic):
                        break
                    c = cols if cols < sc.Nx else 0
                    W = self.window_bins(R, c)
                    if self.lds_bytes(P, nch, self.npol, W) <= lds_budget // 2:
                        return LaunchShape(P, R, W, c)
        for P in (8, 4, 2):
            R = min(8 * (64 // P), sc.Ny)      # 8 waves x 64/P rows
            W = self.window_bins(R)
            if self.lds_bytes(P, nch, self.npol, W) <= lds_budget:
                return LaunchShape(P, R, W)
        raise RuntimeError('range window of %d bins does not fit the LDS even at P=2' % W)

    # ------------------------------------------------------------------
    def new_raw(self, fp64=False):
        """Zeroed raw cube [pol][ch][pulse][bin]: complex64, or complex128 for the
        all-FP64 reference path."""
        import torch
        sc = self.sc
        raw = torch.empty((self.npol, sc.num_ch, self.n_pulses, sc.rg_samp),
                          dtype=torch.complex128 if fp64 else torch.complex64, device=self.dev)
        if fp64:
            self.zero_f64(raw.view(torch.float64))
        else:
            self.zero_c32(raw)
        return raw

    def zero_c32(self, t):
        _lib.check(self.ctx.lib.osar_fill_zero_c32(self.ctx.h, _lib.ptr(t), t.numel(),
                                                   _lib.stream_ptr()))

    def zero_f64(self, t):
        _lib.check(self.ctx.lib.osar_fill_zero_f64(self.ctx.h, _lib.ptr(t), t.numel(),
                                                   _lib.stream_ptr()))

    def bragg(self, diffx, diffy, pulse, amp_out, rcs_sum_out=None, hmtf=None):
        """sqrt(RCS) of every facet at refresh pulse `pulse` into amp_out [npol][2][Nx][Ny]
        (float32; float64 runs the FP64 facet model). hmtf: the snapshot's hMTF
        field, which modulates the RCS (use_hmtf)."""
        import torch
        d = self._bragg_desc()
        d.pulse = int(pulse)
        d.diffx, d.diffy = _lib.ptr(diffx), _lib.ptr(diffy)
        d.hmtf = _lib.ptr(hmtf)
        d.fp64 = 1 if amp_out.dtype == torch.float64 else 0
        if d.fp64:
            d.amp64 = _lib.ptr(amp_out)
        else:
            d.amp = _lib.ptr(amp_out)
        d.rcs_sum = _lib.ptr(rcs_sum_out)
        _lib.check(self.ctx.lib.osar_facet_bragg(self.ctx.h, ctypes.byref(d), _lib.stream_ptr()))

    def bragg_refreshes(self, r_first, r_last, diffx, diffy, snap0, amp_out, rcs_sum_out,
                        hmtf=None):
        """Refreshes r_first..r_last in one launch: slopes from snapshot slots
        (diffx/diffy [slot][Nx][Ny], slot 0 = snapshot snap0), amplitudes into
        amp_out[0 .. r_last - r_first] (float32, or float64 for the FP64 facet
        model), sums into rcs_sum_out[r_first .. r_last]. hmtf: the snapshots'
        hMTF fields (same slots) when the RCS is modulated (use_hmtf)."""
        import torch
        n = r_last - r_first + 1
        d = self._bragg_desc()
        d.fp64 = 1 if amp_out.dtype == torch.float64 else 0
        d.n_refresh = n
        d.snap0 = int(snap0)
        d.refresh_pulse = self.refresh_pulse.data_ptr() + 4 * r_first
        d.refresh_snap = self.refresh_snap.data_ptr() + 4 * r_first
        d.diffx, d.diffy = _lib.ptr(diffx), _lib.ptr(diffy)
        d.hmtf = _lib.ptr(hmtf)
        d.diff_stride = diffx[0].numel()
        if d.fp64:
            d.amp64 = _lib.ptr(amp_out)
        else:
            d.amp = _lib.ptr(amp_out)
        d.amp_stride = amp_out[0].numel()
        d.rcs_sum = rcs_sum_out[r_first].data_ptr()
        _lib.check(self.ctx.lib.osar_facet_bragg(self.ctx.h, ctypes.byref(d), _lib.stream_ptr()))

    def _bragg_desc(self):
        sc, b = self.sc, self.sc.bragg
        d = _lib.BraggDesc()
        d.nx, d.ny = sc.Nx, sc.Ny
        d.pol = POL_CODE[sc.pol] if sc.pol in POL_CODE else 0
        d.spread = b.spread_id
        d.n_spec = len(b.spec_k)
        d.k0, d.d, d.wind_dir, d.wind_U, d.wind_fetch = b.k0, b.d, b.wind_dir, b.wind_U, b.wind_fetch
        d.dx, d.dy = float(sc.surface.dx), float(sc.surface.dy)
        d.t_step, d.t_span, d.v_ground, d.gr0 = sc.t_step, sc.t_span, sc.v_ground, sc.gr0
        d.cos_inc, d.sin_inc = _lib.ptr(self.cos_inc), _lib.ptr(self.sin_inc)
        d.y_row = _lib.ptr(self.y_row)
        d.spec_k, d.spec_S = _lib.ptr(self.spec_k), _lib.ptr(self.spec_S)
        d.spec_slope = _lib.ptr(self.spec_slope)
        d.spec_k0, d.spec_dk, d.spec_uniform = self.spec_grid
        k = b.spec_k
        d.spec_inv_dk = float((len(k) - 1) / (k[-1] - k[0])) if len(k) > 1 else 0.
        return d

    def echo_desc(self, dx, dy, dz, snap0, amp, rcs0, pulse0, npulses, raw, shape=None,
                  spec=None, spec_energy=None):
        """Descriptor of one accumulator launch. spec: dict of slope/curvature snapshot
        tensors (Diffx, Diffy, Diffxx, Diffyy, Diffxy) when specular is enabled."""
        sc = self.sc
        shape = shape or self.shape
        d = _lib.EchoDesc()
        d.pol = POL_CODE.get(sc.pol, 0)
        d.nx, d.ny = sc.Nx, sc.Ny
        d.n_pulses_total = self.n_pulses
        d.pulse0, d.npulses = int(pulse0), int(npulses)
        d.nbins = sc.rg_samp
        d.n_ch, d.n_pol = sc.num_ch, self.npol
        d.K = sc.K
        d.snap0 = int(snap0)
        d.n_snap_slots = int(dx.shape[0])
        d.rcs0 = int(rcs0)
        d.n_rcs_slots = int(amp.shape[0]) if amp is not None else 0
        flags = 0
        if sc.bragg_enable:
            flags |= _lib.ECHO_BRAGG
        if np.any(sc.b_xti != 0):
            flags |= _lib.ECHO_XTI
        if sc.spec_enable:
            if spec is None:
                raise ValueError('specular scattering needs the slope/curvature snapshots')
            flags |= _lib.ECHO_SPECULAR
            d.sx, d.sy = _lib.ptr(spec['Diffx']), _lib.ptr(spec['Diffy'])
            d.sxx, d.syy = _lib.ptr(spec['Diffxx']), _lib.ptr(spec['Diffyy'])
            d.sxy = _lib.ptr(spec['Diffxy'])
            d.inc_col = _lib.ptr(self.inc_col)
            d.facet_dx, d.facet_dy = float(sc.surface.dx), float(sc.surface.dy)
            d.spec_energy = _lib.ptr(spec_energy)
        d.flags = flags
        d.pulses_per_block = shape.pulses_per_block
        d.rows_per_block = shape.rows_per_block
        d.window_bins = shape.window_bins
        d.cols_per_block = shape.cols_per_block
        d.t_step, d.t_span, d.v_ground = sc.t_step, sc.t_span, sc.v_ground
        d.sr0, d.gr0, d.k0 = sc.sr0, sc.gr0, sc.k0
        d.min_sr, d.binsize = sc.min_sr, sc.sr_res / sc.over_fs
        d.ltx_over_l = sc.ant_l_tx / sc.l0
        d.lrx_over_l = sc.ant_l_rx / sc.l0
        d.sr_col_min, d.sr_col_max = float(sc.sr.min()), float(sc.sr.max())
        d.resid_max = self.resid_max
        for i in range(sc.num_ch):
            d.b_ati[i] = sc.b_ati[i]
            d.b_xti[i] = sc.b_xti[i]
        d.sr_col, d.sin_inc, d.cos_inc = (_lib.ptr(self.sr_col), _lib.ptr(self.sin_inc),
                                          _lib.ptr(self.cos_inc))
        d.dinc_col, d.y_row = _lib.ptr(self.dinc_col), _lib.ptr(self.y_row)
        d.dx, d.dy, d.dz = _lib.ptr(dx), _lib.ptr(dy), _lib.ptr(dz)
        d.sinc_tab = _lib.ptr(self.sinc)
        d.raw = _lib.ptr(raw)
        if sc.bragg_enable:
            d.cap_rate = _lib.ptr(self.cap_rate)
            d.amp = _lib.ptr(amp)
            d.rcs_slot = _lib.ptr(self.rcs_slot)
            d.scat_m, d.scat_p = _lib.ptr(self.scat_m), _lib.ptr(self.scat_p)
            d.scat_len = max(self.scat_m.numel(), self.scat_p.numel())
            d.ind0_m, d.ind0_p = _lib.ptr(self.ind0_m), _lib.ptr(self.ind0_p)
        return d

    def accumulate(self, dx, dy, dz, snap0, amp, rcs0, pulse0, npulses, raw, shape=None,
                   spec=None, spec_energy=None):
        """Add pulses [pulse0, pulse0+npulses) of every channel/pol into raw (device).
        A complex128 raw cube (with float64 amplitudes) runs the all-FP64
        reference accumulator osar_echo_accumulate_fp64."""
        import torch
        fp64 = raw.dtype == torch.complex128
        d = self.echo_desc(dx, dy, dz, snap0, amp if not fp64 else None, rcs0, pulse0, npulses,
                           raw, shape, spec, spec_energy)
        if fp64:
            d.raw = None
            d.raw64 = _lib.ptr(raw)
            if self.sc.bragg_enable:
                d.amp = None
                d.amp64 = _lib.ptr(amp)
                d.n_rcs_slots = int(amp.shape[0])
                m, p = self.fp64_tables()
                d.scat64_m, d.scat64_p = _lib.ptr(m), _lib.ptr(p)
            _lib.check(self.ctx.lib.osar_echo_accumulate_fp64(self.ctx.h, ctypes.byref(d),
                                                              _lib.stream_ptr()))
            return
        _lib.check(self.ctx.lib.osar_echo_accumulate(self.ctx.h, ctypes.byref(d),
                                                     _lib.stream_ptr()))

    # ------------------------------------------------------------------
    def fields_needed(self):
        return scene_fields(self.sc)

    def synth(self, times, out, fields=None, slope_fp32=False):
        """Surface fields at `times` into out[name] = float32 [n_times][nx][ny] (x-major)."""
        self._keep = synth_fields(self.ctx, self.sc.surface, self.kx, self.ky, self.coefs,
                                  times, out, fields, slope_fp32=slope_fp32)

    # ------------------------------------------------------------------
    def run(self, p0=0, p1=None, chunk_snaps=32, raw=None, shape=None, timer=None,
            overlap=False, fp64=False):
        """GPU raw cube for pulses [p0, p1): returns (raw device tensor, NRCS_avg numpy).

        Snapshots are synthesised in chunks of `chunk_snaps` intervals, Bragg
        amplitudes of every refresh of the chunk in one launch, then one
        accumulator launch covers the chunk. With `overlap`, synthesis and
        Bragg of chunk c + 1 run on one stream while the accumulation of chunk
        c runs on another (double-buffered snapshots and amplitudes). `timer`,
        if given, is called as timer(stage, start_event, end_event) for every
        launch group. fp64: the all-FP64 reference path (FP64 synthesis
        intermediates, FP64 Bragg facet model, osar_echo_accumulate_fp64) into a
        complex128 raw cube.
        """
        st = self.start(p0, p1, chunk_snaps, raw, shape, overlap=overlap, fp64=fp64)
        while self.step(st, timer):
            pass
        return self.finish(st)

    def start(self, p0=0, p1=None, chunk_snaps=32, raw=None, shape=None, overlap=False,
              fp64=False, nbuf=None):
        """Allocate the buffers of a run over pulses [p0, p1) (no launches).
        nbuf: chunk buffer sets (default 2 with overlap, else 1); with two,
        consecutive prepare() calls alternate between them."""
        import torch
        sc = self.sc
        st = RunState()
        st.p0 = st.p = int(p0)
        st.p1 = self.n_pulses if p1 is None else int(p1)
        st.C = max(1, int(chunk_snaps))
        st.shape = shape
        st.fp64 = bool(fp64)
        if fp64 and overlap:
            raise ValueError('the FP64 reference path runs on one stream')
        st.raw = self.new_raw(fp64) if raw is None else raw
        st.names = self.fields_needed()
        nx, ny = sc.Nx, sc.Ny
        nbuf = (2 if overlap else 1) if nbuf is None else int(nbuf)
        if nbuf not in (1, 2) or (overlap and nbuf != 2):
            raise ValueError('nbuf must be 1 or 2 (2 with overlap)')
        st.bufs = [{n: torch.empty((st.C + 1, nx, ny), dtype=torch.float32, device=self.dev)
                    for n in st.names} for _ in range(nbuf)]
        nref = len(sc.rcs_refresh) if sc.bragg_enable else 0
        st.rcs_sum = torch.zeros(max(nref, 1), self.npol, dtype=torch.float64, device=self.dev)
        st.amps = [None] * nbuf
        if sc.bragg_enable:
            st.amps = [torch.empty((st.C + 2, self.npol, 2, nx, ny),
                                   dtype=torch.float64 if fp64 else torch.float32,
                                   device=self.dev) for _ in range(nbuf)]
        st.k = 0                # buffer set of the next chunk
        st.held = -1            # refresh index whose amplitudes sit in slot held_slot
        st.held_slot = 0
        st.snap_held = -1       # snapshot index held in slot snap_slot
        st.snap_slot = 0
        st.spec_energy = None
        if sc.spec_enable:
            st.spec_energy = torch.zeros(self.n_pulses, dtype=torch.float64, device=self.dev)
        st.streams = None
        if overlap:
            cur = torch.cuda.current_stream(self.dev)
            st.streams = (torch.cuda.Stream(device=self.dev), torch.cuda.Stream(device=self.dev))
            for sx in st.streams:
                sx.wait_stream(cur)
            st.echo_done = [None] * nbuf     # last accumulation that read buffer set k
        return st

    def step(self, st, timer=None):
        """Enqueue the next chunk of `st` (on the current stream, or on the run's
        two streams with overlap); False when done."""
        if st.p >= st.p1:
            return False
        import torch
        if st.streams is None:
            self._step(st, timer, None, None)
            return True
        sa, sb = st.streams
        k = st.k
        with torch.cuda.stream(sa):
            if st.echo_done[k] is not None:
                sa.wait_event(st.echo_done[k])       # the accumulation two chunks back
            ready = self._step(st, timer, sa, sb)
        with torch.cuda.stream(sb):
            done = self._step_echo(st, timer, ready)
        st.echo_done[k] = done
        return True

    def prepare(self, st):
        """Synthesis and Bragg of the next chunk of run `st` on the current
        stream; returns the accumulator descriptor of that chunk, not launched
        (for osar_echo_accumulate_batch), or None when the run is done. The next
        prepare reuses the chunk's buffers: the accumulation must be on the
        stream before it."""
        if st.p >= st.p1:
            return None
        if st.streams is not None or st.fp64:
            raise ValueError('prepare() runs the single-stream mixed-precision path')
        self._step(st, None, None, None, defer=True)
        p, q1, m0, r0 = st.chunk
        buf, amp = st.bufs[st.k], st.amps[st.k]
        d = self.echo_desc(buf['Dx'], buf['Dy'], buf['Dz'], m0, amp, r0, p, q1 - p, st.raw,
                           st.shape, spec=buf if self.sc.spec_enable else None,
                           spec_energy=st.spec_energy)
        if len(st.bufs) > 1:
            st.k = 1 - st.k
        return d

    def _step(self, st, timer, sa, sb, defer=False):
        """Synthesis and Bragg of the next chunk into buffer set st.k (plus the
        accumulation when neither overlapped nor deferred). Returns the event the
        accumulation waits for (overlap) or None."""
        import torch
        sc = self.sc
        K = sc.K
        p, p1, C = st.p, st.p1, st.C
        prev = st.k if len(st.bufs) == 1 else 1 - st.k
        buf, amp = st.bufs[st.k], st.amps[st.k]
        m0 = p // K
        m1 = min(m0 + C, (p1 - 1) // K + 1)       # last snapshot needed
        q1 = min(m1 * K, p1) if m1 > m0 else p1
        if q1 <= p:
            q1 = min(p + K, p1)
        last = q1 - 1
        m_last = last // K + (1 if last % K else 0)
        ev = self._events(timer)
        if st.snap_held == m0 and m_last > m0:
            # the previous chunk's last snapshot is this chunk's first: keep it
            src = st.bufs[prev]
            for n in st.names:
                buf[n][0].copy_(src[n][st.snap_slot])
            self.synth(self.snap_times[m0 + 1:m_last + 1], {n: buf[n][1:] for n in st.names},
                       st.names, self.slope_fp32 and not st.fp64)
        else:
            self.synth(self.snap_times[m0:m_last + 1], buf, st.names,
                       self.slope_fp32 and not st.fp64)
        st.snap_held, st.snap_slot = m_last, m_last - m0
        self._mark(timer, 'synth', ev)
        r0 = r1 = 0
        if sc.bragg_enable:
            r0 = int(sc.rcs_slot[p])
            r1 = int(sc.rcs_slot[last])
            ev = self._events(timer)
            r_first = r0
            if r0 == st.held:
                # the last refresh of the previous chunk is still in force
                if st.held_slot != 0 or prev != st.k:
                    amp[0].copy_(st.amps[prev][st.held_slot])
                r_first = r0 + 1
            if r_first == r0 and int(sc.rcs_refresh[r0]) // K < m0:
                # refresh pulse before this chunk: synthesise its slopes alone
                pr = int(sc.rcs_refresh[r0])
                names = ['Diffx', 'Diffy'] + (['hMTF'] if sc.use_hmtf else [])
                tmp = {n: torch.empty((1, sc.Nx, sc.Ny), dtype=torch.float32, device=self.dev)
                       for n in names}
                self.synth(self.snap_times[pr // K:pr // K + 1], tmp, names,
                           self.slope_fp32 and not st.fp64)
                self.bragg(tmp['Diffx'][0], tmp['Diffy'][0], pr, amp[0], st.rcs_sum[r0],
                           tmp['hMTF'][0] if sc.use_hmtf else None)
                r_first += 1
            if r_first <= r1:
                self.bragg_refreshes(r_first, r1, buf['Diffx'], buf['Diffy'], m0,
                                     amp[r_first - r0:], st.rcs_sum,
                                     buf['hMTF'] if sc.use_hmtf else None)
            self._mark(timer, 'bragg', ev)
            st.held, st.held_slot = r1, r1 - r0
        st.chunk = (p, q1, m0, r0)
        st.p = q1
        if sa is None:
            if not defer:
                self._step_echo(st, timer, None)
            return None
        ready = torch.cuda.Event()
        ready.record(sa)
        return ready

    def _step_echo(self, st, timer, ready):
        """The accumulation of the chunk prepared by _step (buffer set st.k)."""
        import torch
        sc = self.sc
        p, q1, m0, r0 = st.chunk
        buf, amp = st.bufs[st.k], st.amps[st.k]
        if ready is not None:
            torch.cuda.current_stream(self.dev).wait_event(ready)
        ev = self._events(timer)
        self.accumulate(buf['Dx'], buf['Dy'], buf['Dz'], m0, amp, r0, p, q1 - p, st.raw,
                        st.shape, spec=buf if sc.spec_enable else None,
                        spec_energy=st.spec_energy)
        self._mark(timer, 'echo', ev)
        done = None
        if ready is not None:
            done = torch.cuda.Event()
            done.record(torch.cuda.current_stream(self.dev))
        if len(st.bufs) > 1:
            st.k = 1 - st.k
        return done

    def finish(self, st):
        """Wait for the run and return (raw device tensor, NRCS_avg numpy [pol][pulse])."""
        import torch
        sc = self.sc
        if st.streams is not None:
            cur = torch.cuda.current_stream(self.dev)
            for sx in st.streams:
                cur.wait_stream(sx)
        p0, p1 = st.p0, st.p1
        nrcs = np.zeros((self.npol, self.n_pulses))
        if sc.bragg_enable:
            sums = st.rcs_sum.cpu().numpy()
            slots = sc.rcs_slot[p0:p1]
            nrcs[:, p0:p1] = sums[slots].T / sc.surface_area
        if sc.spec_enable:
            # specular |Esn|^2 of the last polarisation goes to every polarisation
            # (sar_raw_nompi.py:427-428)
            nrcs[:, p0:p1] += st.spec_energy[p0:p1].cpu().numpy()[None, :] / sc.surface_area
        return st.raw, nrcs

    @staticmethod
    def _events(timer):
        if timer is None:
            return None
        import torch
        a = torch.cuda.Event(enable_timing=True)
        a.record()
        return a

    @staticmethod
    def _mark(timer, stage, start):
        if timer is None:
            return
        import torch
        b = torch.cuda.Event(enable_timing=True)
        b.record()
        timer(stage, start, b)
