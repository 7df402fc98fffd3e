"""OceanSurface with the reference's interface, synthesised on the GPU.

Mirrors oceansar/surfaces/ocean.py:15-526: init() builds the spectrum and the
random coefficients (host, surface_init.py), load()/save() use the ocean
state file, and setting ``t`` synthesises the requested fields through
osar_surface_synth (FP64 FFTs, float32 results). After ``surface.t = t`` the
fields are available as float32 numpy arrays [Ny, Nx] under the reference's
attribute names (Dx, Dy, Dz, Diffx, Diffy, Diffxx, Diffyy, Diffxy, Vx, Vy, Vz,
Ax, Ay, Az, hMTF) and as device tensors in x-major layout in
``surface.device_fields``.

Supported compute flags: 'D', 'Diff', 'Diff2', 'V', 'A', 'hMTF', choppy or
not; every requested field of one time comes from one osar_surface_synth
call, two fields per FP64 inverse FFT.
"""
import numpy as np

from . import raw_io
from .surface_init import SurfaceInit, init_surface, init_surface_device

_COMPUTE = {'D': ['Dx', 'Dy', 'Dz'], 'Diff': ['Diffx', 'Diffy'],
            'Diff2': ['Diffxx', 'Diffyy', 'Diffxy'], 'V': ['Vx', 'Vy', 'Vz'],
            'A': ['Ax', 'Ay', 'Az'], 'hMTF': ['hMTF']}


class OceanSurface:

    def __init__(self, device=0):
        self.initialized = False
        self.device = device
        self._t = None
        self.device_fields = {}

    # ------------------------------------------------------------------
    def init(self, Lx, Ly, dx, dy, cutoff_wl, spec_model, spread_model, wind_dir, wind_fetch,
             wind_U, current_mag, current_dir, dir_swell_dir=0, freq_r=100, sigf=1, sigs=1,
             Hs=2, swell_dir_enable=False, swell_enable=False, swell_ampl=0., swell_dir=0.,
             swell_wl=0., compute=[], opt_res=True, fft_max_prime=2, choppy_enable=False,
             dirspectrum_func=None, depth=None, dir_swell_spec=None, seed=None, on_device=False):
        """ocean.py:29-242. `seed` (extension) seeds the coefficient draws; None draws
        unseeded, like the reference's global numpy RNG. on_device (extension):
        build the spectrum and the coefficients in HBM (osar_surface_init, Philox
        draws keyed by the seed); wave_coefs is then a device tensor."""
        if dirspectrum_func is not None:
            raise NotImplementedError('external directional spectra (buoy data) are not supported')
        ss = np.random.SeedSequence(seed).spawn(2)
        if on_device:
            s = init_surface_device(Lx, Ly, dx, dy, spec_model, spread_model, wind_dir,
                                    wind_fetch, wind_U, current_mag, current_dir,
                                    seed=int(ss[0].generate_state(1, np.uint64)[0]),
                                    dir_swell_dir=dir_swell_dir, freq_r=freq_r, sigf=sigf,
                                    sigs=sigs, Hs=Hs, swell_dir_enable=swell_dir_enable,
                                    swell_enable=swell_enable, opt_res=opt_res,
                                    fft_max_prime=fft_max_prime, choppy_enable=choppy_enable,
                                    depth=depth, cutoff_wl=cutoff_wl, device=self.device)
            self._adopt(s, compute)
            return
        rngs = [np.random.Generator(np.random.PCG64(s)) for s in ss]
        s = init_surface(Lx, Ly, dx, dy, spec_model, spread_model, wind_dir, wind_fetch, wind_U,
                         current_mag, current_dir, rng_wave=rngs[0], rng_swell=rngs[1],
                         dir_swell_dir=dir_swell_dir, freq_r=freq_r, sigf=sigf, sigs=sigs, Hs=Hs,
                         swell_dir_enable=swell_dir_enable, swell_enable=swell_enable,
                         swell_ampl=swell_ampl, swell_dir=swell_dir, swell_wl=swell_wl,
                         opt_res=opt_res, fft_max_prime=fft_max_prime,
                         choppy_enable=choppy_enable, depth=depth, cutoff_wl=cutoff_wl)
        self._adopt(s, compute)

    def _adopt(self, s, compute):
        self.state = s
        for k in ('Lx', 'Ly', 'Nx', 'Ny', 'dx', 'dy', 'x', 'y', 'wave_coefs', 'rmss_x', 'rmss_y',
                  'wind_U', 'wind_dir', 'wind_U_eff', 'wind_dir_eff', 'wind_fetch', 'current',
                  'swell_enable', 'swell_ampl', 'swell_dir', 'swell_wl', 'swell_ph0'):
            setattr(self, k, getattr(s, k))
        self.choppy_enable = s.choppy
        self.kx, self.ky = np.meshgrid(s.kx_s, s.ky_s)
        self.compute = list(compute)
        for c in self.compute:
            if c not in _COMPUTE:
                raise NotImplementedError("compute '%s' is not available on the GPU path" % c)
        self._dev = None
        self.initialized = True

    # ------------------------------------------------------------------
    def save(self, state_file):
        """ocean.py:334-373."""
        if not self.initialized:
            raise Exception('Surface not initialized')
        s = self.state
        with raw_io.OceanStateFile(state_file, 'w', [s.Ny, s.Nx]) as st:
            st.set('Nx', s.Nx)
            st.set('Ny', s.Ny)
            for k in ('Lx', 'Ly', 'dx', 'dy', 'wind_fetch', 'wind_U', 'wind_U_eff',
                      'swell_ampl', 'swell_wl'):
                st.set(k, float(getattr(s, k)))
            st.set('wind_dir', np.rad2deg(s.wind_dir))
            st.set('wind_dir_eff', np.rad2deg(s.wind_dir_eff))
            cur = np.asarray(s.current, dtype=np.float64)
            st.set('current_mag', float(np.hypot(*cur)))
            st.set('current_dir', float(np.rad2deg(np.arctan2(cur[1], cur[0]))))
            st.set('swell_enable', 1 if s.swell_enable else 0)
            st.set('swell_dir', np.rad2deg(s.swell_dir))
            st.set('swell_ph0', float(s.swell_ph0))
            st.set('choppy_enable', 1 if s.choppy else 0)
            st.set('kx', self.kx)
            st.set('ky', self.ky)
            coefs = s.wave_coefs
            if hasattr(coefs, 'cpu'):   # built on the device
                coefs = coefs.cpu().numpy()
            st.set('wave_coefs*', coefs)
            st._f.depth = -1.0 if s.depth is None else float(s.depth)

    def load(self, state_file, compute=[]):
        """ocean.py:244-332 (the coefficients are read back as complex64)."""
        try:
            st = raw_io.OceanStateFile(state_file, 'r')
        except (OSError, ValueError, TypeError) as e:
            raise RuntimeError('Error loading OTGI File') from e
        with st:
            g = st.get
            Nx, Ny = int(g('Nx')), int(g('Ny'))
            kx = np.asarray(g('kx'), np.float32)
            ky = np.asarray(g('ky'), np.float32)
            coefs = g('wave_coefs*')
            depth = float(getattr(st._f, 'depth', -1.0))
            vals = {k: float(g(k)) for k in raw_io.OceanStateFile.SCALARS}
        kx_s, ky_s = kx[0, :].copy(), ky[:, 0].copy()
        cur_dir = np.deg2rad(vals['current_dir'])
        current = (vals['current_mag'] * np.array([np.cos(cur_dir), np.sin(cur_dir)])
                   ).astype(np.float32)
        s = SurfaceInit(Lx=vals['Lx'], Ly=vals['Ly'], Nx=Nx, Ny=Ny, dx=np.float32(vals['dx']),
                        dy=np.float32(vals['dy']),
                        x=np.linspace(-vals['Lx'] / 2., vals['Lx'] / 2., Nx, dtype=np.float32),
                        y=np.linspace(-vals['Ly'] / 2., vals['Ly'] / 2., Ny, dtype=np.float32),
                        kx_s=kx_s, ky_s=ky_s, kx_res=kx_s[1] - kx_s[0], ky_res=ky_s[1] - ky_s[0],
                        wave_coefs=coefs.astype(np.complex64), wave_dirspec=None,
                        rmss_x=float('nan'), rmss_y=float('nan'), wind_U=vals['wind_U'],
                        wind_dir=np.deg2rad(vals['wind_dir']), wind_U_eff=vals['wind_U_eff'],
                        wind_dir_eff=np.deg2rad(vals['wind_dir_eff']),
                        wind_fetch=vals['wind_fetch'], current=current,
                        depth=None if depth <= 0 else depth,
                        choppy=vals['choppy_enable'] == 1)
        if vals['swell_enable'] == 1:
            raise NotImplementedError('monochromatic swell is not supported on the GPU path')
        self._adopt(s, compute)

    # ------------------------------------------------------------------
    def _device(self):
        if self._dev is None:
            import torch
            from . import _lib
            ctx = _lib.context(self.device)
            dev = torch.device('cuda', self.device)
            s = self.state
            self._dev = (ctx,
                         torch.as_tensor(s.kx_s, device=dev),
                         torch.as_tensor(s.ky_s, device=dev),
                         s.wave_coefs.to(dev) if torch.is_tensor(s.wave_coefs) else
                         torch.as_tensor(np.ascontiguousarray(s.wave_coefs), device=dev))
        return self._dev

    @property
    def t(self):
        if not self.initialized:
            raise Exception('Surface not initialized')
        return self._t

    @t.setter
    def t(self, value):
        """Synthesise every requested field at time `value` (ocean.py:418-526)."""
        if not self.initialized:
            raise Exception('Surface not initialized')
        import torch
        from .engine import synth_fields
        self._t = float(value)
        names = [n for c in self.compute for n in _COMPUTE[c]]
        if not names:
            return
        ctx, kx, ky, coefs = self._device()
        s = self.state
        out = {n: torch.empty((1, s.Nx, s.Ny), dtype=torch.float32, device=coefs.device)
               for n in names}
        keep = synth_fields(ctx, s, kx, ky, coefs, [self._t], out, names)
        torch.cuda.synchronize(coefs.device)
        del keep
        self.device_fields = {n: out[n][0] for n in names}
        for n in names:
            setattr(self, n, np.ascontiguousarray(out[n][0].cpu().numpy().T))
