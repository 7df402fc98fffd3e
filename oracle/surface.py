"""FP64 restatement of the OceanSurface time evolution (oceansar/surfaces/ocean.py:418-526).

Differences from the reference, all towards FP64: t, omega(k), the
dispersion phase, the hMTF multiplier M_h and the inverse FFTs are complex128
(numpy >= 2 runs the reference's ifft2 in complex64; numpy < 2 in
complex128). The hMTF cut k > 2 pi / 20 compares |k| as the reference's
float32 table holds it with 2 pi / 20 rounded to float32 (ocean.py:519-524
compares float32 arrays with a Python float), so the same wave numbers are
cut. Outputs are stored as float32, as OceanSurface stores them
(ocean.py:384-403), or as float64 with dtype=np.float64 (tests only).
"""
import numpy as np

G = 9.80665  # scipy.constants.g


def wave_tables(surf):
    """FP64 k-space tables from the float32 grid (ocean.py:131-158)."""
    kx = surf.kx_s.astype(np.float64)[None, :]
    ky = surf.ky_s.astype(np.float64)[:, None]
    k = np.sqrt(kx ** 2 + ky ** 2)
    good = k > float(np.min(np.array([surf.kx_res, surf.ky_res]))) / 2.0
    ks = np.where(good, k, 1.0)
    kxn = np.where(good, kx / ks, 0.)
    kyn = np.where(good, ky / ks, 0.)
    kinv = np.where(good, 1. / ks, 0.)
    if surf.depth is None:
        omega = np.sqrt(G * k)
    else:
        omega = np.sqrt(G * k * np.tanh(k * float(surf.depth)))
    return kx, ky, k, kxn, kyn, kinv, omega


def _re(a):
    return np.real(np.fft.ifft2(a))


def hmtf_multiplier(surf, k, omega):
    """M_h of ocean.py:511-524 over the grid (FP64), zero where the float32
    |k| exceeds float32(2 pi / 20)."""
    k_max = 2. * np.pi / 20.
    mu = 0.04 / 16.
    dlnS_dlnk = -3.
    Vg_c = 0.5
    M = 1j * k * omega * (omega + 1j * mu) / (omega ** 2. + mu ** 2.) * (dlnS_dlnk - Vg_c)
    kx32 = surf.kx_s.astype(np.float32)[None, :]
    ky32 = surf.ky_s.astype(np.float32)[:, None]
    k32 = np.sqrt(kx32 ** 2 + ky32 ** 2)
    return np.where(k32 > np.float32(k_max), 0., M)


def synth(surf, t, compute=('D', 'Diff'), dtype=np.float32):
    """Fields at time t as a dict of [Ny, Nx] arrays of `dtype` (float32 as
    OceanSurface stores them)."""
    t = float(t)
    kx, ky, k, kxn, kyn, kinv, omega = wave_tables(surf)
    ph = surf.wave_coefs.astype(np.complex128) * np.exp(-1j * omega * t)
    cur = surf.current.astype(np.float64)
    out = {}
    sw = None
    if getattr(surf, 'swell_enable', False):
        # monochromatic swell (ocean.py:218-233, 428-429)
        x = surf.x.astype(np.float64)[None, :]
        y = surf.y.astype(np.float64)[:, None]
        sw_exp = surf.swell_ampl * np.exp(1j * (surf.swell_k * (np.cos(surf.swell_dir) * x +
                                                                np.sin(surf.swell_dir) * y) +
                                                surf.swell_ph0))
        sw = sw_exp * np.exp(-1j * surf.swell_omega * t)
        skx, sky, sk = surf.swell_kx, surf.swell_ky, surf.swell_k
    if 'D' in compute:
        Dx = -np.imag(np.fft.ifft2(kxn * ph)) + cur[0] * t
        Dy = -np.imag(np.fft.ifft2(kyn * ph)) + cur[1] * t
        Dz = _re(ph)
        if sw is not None:
            Dx = Dx + np.real(1j * skx / sk * sw)
            Dy = Dy + np.real(1j * sky / sk * sw)
            Dz = Dz + np.real(sw)
        out.update(Dx=Dx, Dy=Dy, Dz=Dz)
    if 'Diff' in compute:
        if not surf.choppy:
            Diffx = _re(1j * kx * ph)
            Diffy = _re(1j * ky * ph)
        else:
            Diffx = _re(1j * kx * ph) / (1. + _re(-kx ** 2. * kinv * ph))
            Diffy = _re(1j * ky * ph) / (1. + _re(-ky ** 2. * kinv * ph))
        if sw is not None:
            Diffx = Diffx + np.real(1j * skx * sw)
            Diffy = Diffy + np.real(1j * sky * sw)
        out.update(Diffx=Diffx, Diffy=Diffy)
    if 'Diff2' in compute:
        if not surf.choppy:
            Dxx = _re(-kx ** 2. * ph)
            Dyy = _re(-ky ** 2. * ph)
            Dxy = _re(-kx * ky * ph)
        else:
            a_x = _re(1j * kx * ph)
            a_y = _re(1j * ky * ph)
            a_xx = _re(-kx ** 2. * kinv * ph)
            a_yy = _re(-ky ** 2. * kinv * ph)
            a_xxx = _re(-1j * kx ** 3. * kinv * ph)
            a_yyy = _re(-1j * ky ** 3. * kinv * ph)
            a_xxy = _re(-1j * kx ** 2. * ky * kinv * ph)
            Dxx = ((1. + a_xx) * _re(-kx ** 2. * ph) - a_xxx * a_x) / ((1 + a_xx) ** 3.)
            Dyy = ((1. + a_yy) * _re(-ky ** 2. * ph) - a_yyy * a_y) / ((1 + a_yy) ** 3.)
            Dxy = (((1. + a_xx) * _re(-kx * ky * ph) - a_xxy * a_x) /
                   ((1 + a_xx) ** 2. * (1 + a_yy)))
        if sw is not None:
            Dxx = Dxx + np.real(-skx ** 2. * sw)
            Dyy = Dyy + np.real(-sky ** 2. * sw)
            Dxy = Dxy + np.real(-skx * sky * sw)
        out.update(Diffxx=Dxx, Diffyy=Dyy, Diffxy=Dxy)
    if 'V' in compute:
        v = -1j * omega * ph
        Vx = -np.imag(np.fft.ifft2(kxn * v)) + cur[0]
        Vy = -np.imag(np.fft.ifft2(kyn * v)) + cur[1]
        Vz = _re(v)
        if sw is not None:
            sv = -1j * surf.swell_omega * sw
            Vx = Vx + np.real(1j * skx / sk * sv)
            Vy = Vy + np.real(1j * sky / sk * sv)
            Vz = Vz + np.real(sv)
        out.update(Vx=Vx, Vy=Vy, Vz=Vz)
    if 'A' in compute:
        a = -omega ** 2. * ph
        Ax = _re(1j * kx * kinv * a)
        Ay = _re(1j * ky * kinv * a)
        Az = _re(a)
        if sw is not None:
            sa = -surf.swell_omega ** 2. * sw
            Ax = Ax + np.real(1j * skx / sk * sa)
            Ay = Ay + np.real(1j * sky / sk * sa)
            Az = Az + np.real(sa)
        out.update(Ax=Ax, Ay=Ay, Az=Az)
    if 'hMTF' in compute:
        # monochromatic swell does not enter (ocean.py:508-526)
        out['hMTF'] = _re(hmtf_multiplier(surf, k, omega) * ph)
    return {n: np.ascontiguousarray(v, dtype=dtype) for n, v in out.items()}
