"""GPU parity: InSAR L1b and ATI processing (osar_insar_coregister,
osar_insar_multilook, insar_process, ati_process) against the numpy oracle
(oracle/insar.py).

Bounds: the shifted SLC is rounded to complex64 on both sides (the reference
stores it into its complex64 array), so it agrees to within one float32
rounding of the FP64 shift (relative RMS below 1e-7, nearly every value
identical); the multilooked products are FP64 on both sides: flat windows sum
in the reference's order (relative error below 1e-13), tap windows differ from
scipy's fftconvolve by its FFT rounding (below 1e-12 of the image's peak);
the coherence phase to 1e-11 rad.
"""
import numpy as np
import pytest

import oracle.insar as oi

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _ctx():
    from oceansar_amd import _lib
    return _lib.context(0)


@pytest.mark.parametrize('naz,nrg,npol,flatten', [
    (1, 3, 1, False), (37, 5, 1, True), (64, 40, 2, False), (1000, 115, 1, True),
    (4095, 33, 1, False), (4097, 24, 2, True), (7555, 624, 1, True)])
def test_coregister_matches_oracle(need_gpu, naz, nrg, npol, flatten):
    """Lines of any length: one LDS transform per column up to 4096 lines
    (M <= 8192), the two-half form beyond (M = 16384, up to 8192 lines); 7555
    lines x 624 samples is the large workload's SLC."""
    import torch
    from oceansar_amd.insar_processor import CoregPlan, coregister_channel
    rng = np.random.default_rng(naz + nrg)
    x = (rng.standard_normal((1, npol, naz, nrg)) +
         1j * rng.standard_normal((1, npol, naz, nrg))).astype(np.complex64)
    b, v, prf = 7.5, 6900., 1800.
    flat = np.exp(1j * rng.uniform(-3, 3, nrg)) if flatten else None
    dev = torch.as_tensor(x[0], device='cuda')
    got = coregister_channel(_ctx(), dev, CoregPlan(naz, -b / (v / prf), flat)).cpu().numpy()
    want = oi.coregister(x.copy(), [b], v, prf)[0].astype(np.complex128)
    if flatten:
        want = want * flat[None, None, :]
    e = rel_rms(got, want)
    same = float(np.mean(got == want))
    print('coregister %d x %d x %d flatten=%s: rel RMS %.2e, identical %.4f'
          % (npol, naz, nrg, flatten, e, same))
    assert e < 1e-7
    assert np.max(np.abs(got - want)) <= 2.5e-7 * np.max(np.abs(want))
    if not flatten:
        assert same > 0.98


@pytest.mark.parametrize('window,L', [('flat', 1), ('flat', 2), ('flat', 5), ('flat', 8),
                                      ('hanning', 2), ('hanning', 5), ('hamming', 8),
                                      ('bartlett', 6), ('blackman', 7)])
def test_multilook_matches_oracle(need_gpu, window, L):
    import torch
    from oceansar_amd.insar_processor import multilook
    rng = np.random.default_rng(L)
    C, naz, nrg = 4, 96, 70
    x = rng.standard_normal((C, 1, naz, nrg)) + 1j * rng.standard_normal((C, 1, naz, nrg))
    crop = (3, 90, 0, 64)
    i_all, cohs, lut = oi.l1b(x, L, L + 1, window, *crop)
    inten, coh, pha = multilook(_ctx(), torch.as_tensor(x[:, 0], device='cuda'), L, L + 1,
                                window, crop)
    inten, coh, pha = inten.cpu().numpy(), coh.cpu().numpy(), pha.cpu().numpy()
    # scipy's fftconvolve errs by FFT rounding relative to the image's peak, so tap
    # windows are compared against the peak; flat windows value by value
    scale = np.abs(i_all[:, 0]) if window == 'flat' else np.max(np.abs(i_all))
    e_i = np.max(np.abs(inten - i_all[:, 0]) / scale)
    e_c = np.max(np.abs(coh - np.abs(cohs)))
    e_p = np.max(np.abs(np.angle(np.exp(1j * (pha - np.angle(cohs))))))
    print('multilook %s L=%d: intensity max rel %.2e, |coh| max %.2e, phase max %.2e rad'
          % (window, L, e_i, e_c, e_p))
    tol = 1e-13 if window == 'flat' else 1e-12
    assert e_i < tol and e_c < 10 * tol and e_p < 1e-11


def test_multilook_dual_pol_edges_and_real_images(need_gpu):
    """Two channels x two polarisations (6 pairs, the reference's LUT order),
    a crop touching every edge; real images smoothed as they are."""
    import torch
    from oceansar_amd.insar_processor import multilook
    rng = np.random.default_rng(7)
    x = rng.standard_normal((2, 2, 50, 33)) + 1j * rng.standard_normal((2, 2, 50, 33))
    i_all, cohs, lut = oi.l1b(x, 4, 3, 'flat', 0, 50, 0, 33)
    inten, coh, pha = multilook(_ctx(), torch.as_tensor(x.reshape(4, 50, 33), device='cuda'),
                                4, 3, 'flat')
    np.testing.assert_allclose(inten.cpu().numpy().reshape(2, 2, 50, 33), i_all, rtol=1e-13)
    got = coh.cpu().numpy() * np.exp(1j * pha.cpu().numpy())
    np.testing.assert_allclose(got, cohs, rtol=0, atol=1e-12)
    f = rng.standard_normal((61, 47))
    want = oi.smooth(oi.smooth(f, 9, axis=1), 4, axis=0)
    out, c2, p2 = multilook(_ctx(), torch.as_tensor(f[None], device='cuda'), 9, 4, 'flat')
    assert c2 is None and p2 is None
    np.testing.assert_array_equal(out.cpu().numpy()[0], want)


def test_insar_and_ati_end_to_end(need_gpu, tmp_path):
    """sarsim with every stage on a 2-channel ATI scene: sar_raw, sar_focus,
    insar_process and ati_process on the GPU. The L1b file equals the oracle's
    processing of the same SLC file, and the ATI report's numbers the oracle's
    statistics of the same L1b file and ocean state."""
    import oracle.surface as osurf
    from oceansar_amd import configs, raw_io
    from oceansar_amd.geometry import inc_to_sr, sr_to_geo
    from oceansar_amd.ocean import OceanSurface
    from oceansar_amd.sar_raw import sarsim
    text = configs.workload('small', num_ch=2, b_ati=[0., 10.], stages=('proc', 'insar', 'ati'))
    text = text.replace('[sim]\n', '[sim]\npath=%s\n' % tmp_path, 1)
    cfg = configs.write(str(tmp_path / 'sim.cfg'), text)
    sarsim(cfg)
    # oracle: the SLC file through the reference's InSAR steps
    with raw_io.ProcFile(str(tmp_path / 'proc_data.nc'), 'r') as pf:
        proc = pf.get('slc*')
        sr0, prf, v, f0 = pf.get('sr0'), pf.get('prf'), pf.get('v_ground'), pf.get('f0')
        b_ati, b_xti = pf.get('b_ati'), pf.get('b_xti')
        rg_sampling, alt = pf.get('rg_sampling'), pf.get('orbit_alt')
        inc = np.deg2rad(pf.get('inc_angle'))
    surf = OceanSurface()
    surf.load(str(tmp_path / 'ocean.nc'))
    ref_f = osurf.synth(surf.state, 0., ('D', 'V'))
    _, _, _, avg, std = oi.velocity_stats(ref_f['Vx'], ref_f['Vz'], inc, v, sr0)
    az_min, az_max, rg_min, rg_max, _, ok = oi.crop(proc.shape[2], surf.Lx, surf.Ly, rg_sampling,
                                                   inc, v, prf, avg, std)
    assert ok
    p = oi.coregister(proc.copy(), b_ati, v, prf)
    k0 = 2 * np.pi * f0 / oi.C
    p = p * oi.flattening_phasor(2, proc.shape[3], sr0, rg_sampling, alt, k0, b_xti, inc,
                                 sr_to_geo)
    i_all, cohs, lut = oi.l1b(p, 5, 5, 'hanning', az_min, az_max, rg_min, rg_max)
    with raw_io.L1bFile(str(tmp_path / 'insar_l1b.nc'), 'r') as lf:
        gi = lf.get('ml_intensity')
        gc = lf.get('ml_coherence') * np.exp(1j * lf.get('ml_phase'))
        glut = lf.get('coh_lut')
        g_inc = lf.get('inc_angle')
    e_i, e_c = rel_rms(gi, i_all), float(np.max(np.abs(gc - cohs)))
    print('L1b %s: intensity rel RMS %.2e, coherence max abs %.2e; mean |coh| %.3f'
          % (gi.shape, e_i, e_c, np.mean(np.abs(cohs))))
    assert gi.shape == i_all.shape and e_i < 1e-6 and e_c < 1e-5
    assert glut[0, 0, 1, 0] == 0 and g_inc == pytest.approx(inc)
    # ATI report: the L1b file's values (inc_angle converted once more, as the reference reads it)
    a_inc = np.deg2rad(g_inc)
    want = oi.ati_stats(gi, gc, glut, b_ati, f0, v, prf, a_inc, rg_sampling, 5, 5, 2,
                        ref_f['Vx'], ref_f['Vz'], surf.dx, surf.dy, inc_to_sr(a_inc, alt))
    report = open(str(tmp_path / 'ati_data.txt')).read()
    assert 'ESTIMATED RADIAL VELOCITY - NO SMOOTHING' in report
    import re
    from oceansar_amd.ati_processor import ati_process
    got = ati_process(cfg, str(tmp_path / 'insar_l1b.nc'), str(tmp_path / 'ocean.nc'),
                      str(tmp_path / 'ati2.txt'))
    for k in ('v_radial_surf_mean', 'v_radial_surf_std', 'v_radial_surf_ml_std',
              'v_horizo_surf_mean', 'v_horizo_surf_std', 'v_horizo_surf_ml_std',
              'v_radial_est_mean', 'v_radial_est_std', 'v_radial_est_ml_mean',
              'v_radial_est_ml_std', 'nrcs', 'v_r_dop', 'coh_mean', 'abscoh_mean'):
        print('ati %s: %s vs %s' % (k, np.ravel(got[k])[:2], np.ravel(want[k])[:2]))
        np.testing.assert_allclose(np.asarray(got[k], dtype=complex),
                                   np.asarray(want[k], dtype=complex), rtol=2e-5, atol=1e-6)
    assert got['az_guard'] == want['az_guard']
    for line in ('SURFACE RADIAL VELOCITY - SMOOTHING (WIN. SIZE=5x5)', 'vv Polarization',
                 'Estimated mean NRCS', 'GROUND RANGE GRID SPACING', 'AZIMUTH GRID SPACING'):
        assert line in report
    nrcs_line = re.search(r'Estimated mean NRCS = *(-?[0-9.]+)', report)
    assert float(nrcs_line.group(1)) == pytest.approx(float(want['nrcs'][0, 0]), abs=0.006)
