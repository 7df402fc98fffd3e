"""GPU parity: FP64 surface synthesis and the end-to-end GPU pulse loop vs the FP64 oracle.

Synthesis: the device fields are the FP64 inverse FFTs rounded to float32;
the oracle's are numpy complex128 ifft2 rounded to float32. Both sides round
the same FP64 value unless it lies within ~1e-15 relative of a float32
rounding boundary, so we require >= 99.9 % bitwise-equal values and the rest
within one float32 ulp-scale of the field RMS.
End to end: raw within 1e-4 relative RMS and 1 mrad phase RMS of the FP64
oracle (the parity bound of the path), on windows that cross synthesis chunks and RCS
refreshes.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def phase_rms(a, b, floor_db=-30.):
    mag = np.abs(b)
    keep = mag > mag.max() * 10 ** (floor_db / 20.)
    d = np.angle(a[keep] * np.conj(b[keep]))
    return float(np.sqrt(np.mean(d ** 2)))


def _scene(tmp_path_factory, name, **kw):
    from oceansar_amd import configs
    base = kw.pop('base', 'small')
    return configs.load_scene(configs.workload(base, **kw), tmp_path_factory.mktemp(name))


def _synth_compare(sc, times, names, compute):
    import torch
    import oracle.surface as osurf
    from oceansar_amd.engine import DeviceScene
    ds = DeviceScene(sc)
    out = {n: torch.empty((len(times), sc.Nx, sc.Ny), dtype=torch.float32, device='cuda')
           for n in names}
    ds.synth(times, out, names)
    torch.cuda.synchronize()
    worst = 0.
    for ti, t in enumerate(times):
        ref = osurf.synth(sc.surface, t, compute)
        for n in names:
            got = out[n][ti].cpu().numpy().T
            r = ref[n]
            if not np.any(r):
                # a field that vanishes (hMTF without waves longer than 20 m) keeps
                # the FP64 rounding its spectrum's partner field leaves in it
                assert np.max(np.abs(got)) < 1e-12, (n, t, np.max(np.abs(got)))
                continue
            same = np.mean(got == r)
            err = np.max(np.abs(got - r)) / (np.sqrt(np.mean(r.astype(np.float64) ** 2)) + 1e-30)
            worst = max(worst, err)
            assert same > 0.999, (n, t, same)
            assert err < 1e-6, (n, t, err)
    return worst


def test_synth_slopes_in_complex64_workspace(need_gpu, tmp_path_factory):
    """slope_fp32 (the Bragg-only run loop): the displacements stay FP64 end to
    end (bit-equal to the oracle's rounding), the slopes carry the float32
    rounding of the row-pass result only (max error well below 1e-6 of their RMS)."""
    import torch
    import oracle.surface as osurf
    from oceansar_amd.engine import DeviceScene
    sc = _scene(tmp_path_factory, 'sf32', L=512., dx=1.)
    ds = DeviceScene(sc)
    assert ds.slope_fp32
    names = ['Dx', 'Dy', 'Dz', 'Diffx', 'Diffy']
    times = [0.0, 3.3, 17.25]
    out = {n: torch.empty((len(times), sc.Nx, sc.Ny), dtype=torch.float32, device='cuda')
           for n in names}
    ds.synth(times, out, names, slope_fp32=True)
    torch.cuda.synchronize()
    for ti, t in enumerate(times):
        ref = osurf.synth(sc.surface, t, ('D', 'Diff'))
        for n in names:
            got = out[n][ti].cpu().numpy().T
            r = ref[n]
            same = np.mean(got == r)
            err = np.max(np.abs(got - r)) / np.sqrt(np.mean(r.astype(np.float64) ** 2))
            print('%s t=%g: %.4f bit-equal, max err %.2e RMS' % (n, t, same, err))
            if n.startswith('D') and not n.startswith('Diff'):
                assert same > 0.999, (n, t, same)
                assert err < 1e-6, (n, t, err)
            else:
                assert err < 5e-7, (n, t, err)


def test_synth_matches_oracle_square(need_gpu, tmp_path_factory):
    sc = _scene(tmp_path_factory, 's1')
    names = ['Dx', 'Dy', 'Dz', 'Diffx', 'Diffy', 'Diffxx', 'Diffyy', 'Diffxy']
    _synth_compare(sc, [0.0, 0.3711, 4.25], names, ('D', 'Diff', 'Diff2'))


def test_synth_matches_oracle_rectangular_choppy(need_gpu, tmp_path_factory):
    """Nx != Ny, current, finite depth and choppy slopes."""
    from oceansar_amd import configs
    text = configs.workload('small', L=128., dx=1., choppy=True)
    text = text.replace('Ly=128.0', 'Ly=64.0').replace('current_mag=0', 'current_mag=0.4')
    text = text.replace('depth=1000', 'depth=20')
    sc = configs.load_scene(text, tmp_path_factory.mktemp('s2'))
    assert (sc.Nx, sc.Ny) == (128, 64)
    _synth_compare(sc, [0.0, 1.7], ['Dx', 'Dy', 'Dz', 'Diffx', 'Diffy'], ('D', 'Diff'))


ALL_FIELDS = ['Dx', 'Dy', 'Dz', 'Diffx', 'Diffy', 'Diffxx', 'Diffyy', 'Diffxy', 'Vx', 'Vy',
              'Vz', 'Ax', 'Ay', 'Az', 'hMTF']


def test_synth_all_fields_match_oracle(need_gpu, tmp_path_factory):
    """Every OceanSurface field in one call: the 14 monomial fields pair into 7
    spectra, the hydrodynamic MTF (a complex multiplier even in k) takes an
    eighth."""
    sc = _scene(tmp_path_factory, 'sall')
    _synth_compare(sc, [0.0, 0.3711, 4.25], ALL_FIELDS, ('D', 'Diff', 'Diff2', 'V', 'A', 'hMTF'))


def test_synth_time_derivatives_rectangular_choppy_current(need_gpu, tmp_path_factory):
    """Velocities, accelerations and hMTF next to choppy slopes, Nx != Ny, a
    current (Vx, Vy carry it) and finite depth."""
    from oceansar_amd import configs
    text = configs.workload('small', L=256., dx=1., choppy=True)
    text = text.replace('Ly=256.0', 'Ly=128.0').replace('current_mag=0', 'current_mag=0.4')
    text = text.replace('current_dir=90.', 'current_dir=30.').replace('depth=1000', 'depth=20')
    sc = configs.load_scene(text, tmp_path_factory.mktemp('sva'))
    assert (sc.Nx, sc.Ny) == (256, 128) and sc.surface.choppy
    names = ['Dx', 'Dy', 'Dz', 'Diffx', 'Diffy', 'Vx', 'Vy', 'Vz', 'Ax', 'Ay', 'Az', 'hMTF']
    _synth_compare(sc, [0.0, 2.9], names, ('D', 'Diff', 'V', 'A', 'hMTF'))


@pytest.mark.parametrize('lx', [128., 16.])
def test_synth_choppy_curvatures(need_gpu, tmp_path_factory, lx):
    """Choppy second derivatives (ocean.py:462-475): five FP64 spectra kept in the
    workspace by the column pass and combined per point, beside every other
    field (ten spectra, the most a call takes); 16-point transforms too."""
    from oceansar_amd import configs
    text = configs.workload('small', L=lx, dx=1., choppy=True)
    text = text.replace('current_mag=0', 'current_mag=0.3')
    sc = configs.load_scene(text, tmp_path_factory.mktemp('sc2'))
    assert sc.surface.choppy
    _synth_compare(sc, [0.0, 1.9, 6.5], ALL_FIELDS, ('D', 'Diff', 'Diff2', 'V', 'A', 'hMTF'))
    _synth_compare(sc, [2.2], ['Diffxx', 'Diffxy'], ('Diff2',))


@pytest.mark.parametrize('lx,ly', [(16., 32.), (32., 16.)])
def test_synth_matches_oracle_16_point_transforms(need_gpu, tmp_path_factory, lx, ly):
    """16-point row or column transforms: the padded-address fallback of the
    LDS FFT (longer transforms take the short addressing)."""
    from oceansar_amd import configs
    text = configs.workload('small', L=lx, dx=1.)
    text = text.replace('Ly=%.1f' % lx, 'Ly=%.1f' % ly)
    sc = configs.load_scene(text, tmp_path_factory.mktemp('s16'))
    assert (sc.Nx, sc.Ny) == (int(lx), int(ly))
    names = ['Dx', 'Dy', 'Dz', 'Diffx', 'Diffy', 'Diffxx', 'Diffyy', 'Diffxy']
    _synth_compare(sc, [0.0, 2.5], names, ('D', 'Diff', 'Diff2'))


@pytest.mark.parametrize('lx', [16., 256.])
def test_synth_general_row_pass_matches_mirror_pass(need_gpu, tmp_path_factory, lx):
    """The row pass for arbitrary wave-number tables (mirror_grid = 0: a(k) of
    the Nyquist row and column recomputed where needed) and the one for the
    2 pi fftfreq grid (kept in LDS) give the same bits."""
    import torch
    from oceansar_amd import configs
    from oceansar_amd.engine import DeviceScene, synth_fields
    sc = configs.load_scene(configs.workload('small', L=lx, dx=1.), tmp_path_factory.mktemp('sg'))
    ds = DeviceScene(sc)
    names = ALL_FIELDS
    res = []
    for mirror in (True, False):
        out = {n: torch.empty((2, sc.Nx, sc.Ny), dtype=torch.float32, device='cuda')
               for n in names}
        keep = synth_fields(ds.ctx, sc.surface, ds.kx, ds.ky, ds.coefs, [0.7, 3.1], out, names,
                            mirror=mirror)
        torch.cuda.synchronize()
        del keep
        res.append({n: out[n].cpu().numpy() for n in names})
    for n in names:
        assert np.array_equal(res[0][n], res[1][n]), n


@pytest.mark.parametrize('lx,times', [(16., [0.3, 2.9, 7.7]), (256., [0.7, 3.1, 5.2, 11.0]),
                                      (256., [4.4])])
def test_synth_time_paired_field(need_gpu, tmp_path_factory, lx, times):
    """The five fields of the Bragg-only run loop: (Dz, Dx) pair up, the slopes
    share a complex64 spectrum, and Dy pairs with itself at the next snapshot
    (one transform for snapshots ta and ta + 1; a last odd snapshot alone, a
    single snapshot without pairing). The general row pass (any wave-number
    tables) and the mirror-grid pass give the same bits, every field matches
    the oracle as the unpaired synthesis does, and a second call gives the
    same bits (a self-paired row stores each column once)."""
    import torch
    import oracle.surface as osurf
    from oceansar_amd import configs
    from oceansar_amd.engine import DeviceScene, synth_fields
    sc = configs.load_scene(configs.workload('small', L=lx, dx=1.), tmp_path_factory.mktemp('tp'))
    ds = DeviceScene(sc)
    names = ['Dx', 'Dy', 'Dz', 'Diffx', 'Diffy']
    res = []
    for mirror in (True, False, True):
        out = {n: torch.empty((len(times), sc.Nx, sc.Ny), dtype=torch.float32, device='cuda')
               for n in names}
        keep = synth_fields(ds.ctx, sc.surface, ds.kx, ds.ky, ds.coefs, times, out, names,
                            mirror=mirror, slope_fp32=True)
        torch.cuda.synchronize()
        del keep
        res.append({n: out[n].cpu().numpy() for n in names})
    for n in names:
        assert np.array_equal(res[0][n], res[1][n]), ('general vs mirror', n)
        assert np.array_equal(res[0][n], res[2][n]), ('repeat', n)
    for ti, t in enumerate(times):
        ref = osurf.synth(sc.surface, t, ('D', 'Diff'))
        for n in names:
            got, r = res[0][n][ti].T, ref[n]
            err = np.max(np.abs(got - r)) / np.sqrt(np.mean(r.astype(np.float64) ** 2))
            if n in ('Dx', 'Dy', 'Dz'):
                assert np.mean(got == r) > 0.999, (n, t)
                assert err < 1e-6, (n, t, err)
            else:
                assert err < 5e-7, (n, t, err)


def test_end_to_end_matches_oracle(need_gpu, tmp_path_factory):
    """Full GPU loop (synthesis, Bragg, echo) on a pulse window crossing chunks."""
    import torch
    import oracle.sar as osar
    from oceansar_amd.engine import DeviceScene
    sc = _scene(tmp_path_factory, 'e1')
    ds = DeviceScene(sc)
    p0, p1 = 450, 590
    raw, nrcs = ds.run(p0, p1, chunk_snaps=3)
    torch.cuda.synchronize()
    got = raw[:, :, p0:p1].cpu().numpy().astype(np.complex128)
    ref, ref_nrcs = osar.simulate(sc, p0, p1)
    e, ph = rel_rms(got, ref), phase_rms(got, ref)
    print('end-to-end rel RMS %.3e phase RMS %.3e' % (e, ph))
    assert e < 1e-4 and ph < 1e-3
    # the Bragg facet model runs in float32 and Bragg-only runs take the slopes
    # through a complex64 workspace: the RCS sums agree to ~1e-7
    print('NRCS max rel diff %.2e' % np.max(np.abs(nrcs[:, p0:p1] / ref_nrcs - 1)))
    np.testing.assert_allclose(nrcs[:, p0:p1], ref_nrcs, rtol=1e-6)


@pytest.mark.parametrize('K', [1, 3])
def test_end_to_end_snapshot_cadence(need_gpu, tmp_path_factory, K):
    """K = 1 is the reference's own cadence (a surface and an RCS refresh at
    every pulse, no interpolation); K = 3 does not divide the 8 pulses of a
    workgroup, so its lanes span four snapshots. Both against the FP64 oracle
    on a window crossing launch groups."""
    import torch
    import oracle.sar as osar
    from oceansar_amd.engine import DeviceScene
    sc = _scene(tmp_path_factory, 'k%d' % K, L=64., n_pulses=96, K=K, num_ch=2,
                b_ati=[0., 10.])
    ds = DeviceScene(sc)
    p0, p1 = 29, 71
    raw, nrcs = ds.run(p0, p1, chunk_snaps=max(1, 16 // K))
    torch.cuda.synchronize()
    got = raw[:, :, p0:p1].cpu().numpy().astype(np.complex128)
    ref, ref_nrcs = osar.simulate(sc, p0, p1)
    e, ph = rel_rms(got, ref), phase_rms(got, ref)
    print('K=%d rel RMS %.3e phase RMS %.3e' % (K, e, ph))
    assert e < 1e-4 and ph < 1e-3
    np.testing.assert_allclose(nrcs[:, p0:p1], ref_nrcs, rtol=1e-6)


def test_end_to_end_dp_multichannel_xti(need_gpu, tmp_path_factory):
    """DP polarisation (the hh mask quirk), 3 channels with ATI and XTI baselines."""
    import torch
    import oracle.sar as osar
    from oceansar_amd.engine import DeviceScene
    sc = _scene(tmp_path_factory, 'e2', pol='DP', num_ch=3, b_ati=[0., 7., 15.],
                b_xti=[0., 30., -12.], spread='elfouhaily', K=4, wind_dir=30.)
    ds = DeviceScene(sc)
    p0, p1 = 500, 530
    raw, nrcs = ds.run(p0, p1, chunk_snaps=2)
    torch.cuda.synchronize()
    got = raw[:, :, p0:p1].cpu().numpy().astype(np.complex128)
    ref, ref_nrcs = osar.simulate(sc, p0, p1)
    assert got.shape == ref.shape == (2, 3, p1 - p0, sc.rg_samp)
    e, ph = rel_rms(got, ref), phase_rms(got, ref)
    print('DP/XTI rel RMS %.3e phase RMS %.3e' % (e, ph))
    assert e < 1e-4 and ph < 1e-3
    # the Bragg facet model runs in float32 and Bragg-only runs take the slopes
    # through a complex64 workspace: the RCS sums agree to ~1e-7
    print('NRCS max rel diff %.2e' % np.max(np.abs(nrcs[:, p0:p1] / ref_nrcs - 1)))
    np.testing.assert_allclose(nrcs[:, p0:p1], ref_nrcs, rtol=1e-6)


@pytest.mark.parametrize('pol', ['vv', 'DP'])
def test_end_to_end_specular(need_gpu, tmp_path_factory, pol):
    """Bragg + Kirchhoff stationary-phase specular (ati_swell workload physics at 8 deg incidence,
    where stationary-phase facets exist), two ATI channels."""
    import torch
    import oracle.sar as osar
    from oceansar_amd.engine import DeviceScene
    sc = _scene(tmp_path_factory, 'k' + pol, base='ati_swell', L=128., n_pulses=512, inc=8., pol=pol)
    assert sc.spec_enable
    ds = DeviceScene(sc)
    p0, p1 = 250, 262
    raw, nrcs = ds.run(p0, p1, chunk_snaps=1)
    torch.cuda.synchronize()
    got = raw[:, :, p0:p1].cpu().numpy().astype(np.complex128)
    ref, ref_nrcs = osar.simulate(sc, p0, p1)
    e, ph = rel_rms(got, ref), phase_rms(got, ref)
    print('specular %s rel RMS %.3e phase RMS %.3e' % (pol, e, ph))
    assert e < 1e-4 and ph < 1e-3
    # the Bragg facet model runs in float32: its RCS sums agree to ~1e-7
    print('NRCS max rel diff %.2e' % np.max(np.abs(nrcs[:, p0:p1] / ref_nrcs - 1)))
    np.testing.assert_allclose(nrcs[:, p0:p1], ref_nrcs, rtol=1e-6)


def test_end_to_end_specular_choppy(need_gpu, tmp_path_factory):
    """Specular scattering on a choppy surface: the stationary-phase facets
    read the choppy slopes and curvatures."""
    import torch
    import oracle.sar as osar
    from oceansar_amd.engine import DeviceScene
    sc = _scene(tmp_path_factory, 'kc', base='ati_swell', L=128., n_pulses=512, inc=8.,
                choppy=True)
    assert sc.spec_enable and sc.surface.choppy
    ds = DeviceScene(sc)
    p0, p1 = 250, 262
    raw, nrcs = ds.run(p0, p1, chunk_snaps=1)
    torch.cuda.synchronize()
    got = raw[:, :, p0:p1].cpu().numpy().astype(np.complex128)
    ref, ref_nrcs = osar.simulate(sc, p0, p1)
    e, ph = rel_rms(got, ref), phase_rms(got, ref)
    print('specular choppy rel RMS %.3e phase RMS %.3e' % (e, ph))
    assert e < 1e-4 and ph < 1e-3
    np.testing.assert_allclose(nrcs[:, p0:p1], ref_nrcs, rtol=1e-6)


@pytest.mark.parametrize('pol', ['vv', 'DP'])
def test_end_to_end_hmtf(need_gpu, tmp_path_factory, pol):
    """use_hmtf (sar_raw_nompi.py:448-456): the Bragg RCS of every refresh times
    1 + hMTF of its snapshot (hMTF synthesised beside the slopes, paired with
    the odd displacement field in one spectrum), on a window crossing chunks and
    refreshes; the modulation moves the NRCS by more than the parity bound."""
    import torch
    import oracle.sar as osar
    from oceansar_amd.engine import DeviceScene
    sc = _scene(tmp_path_factory, 'hm' + pol, pol=pol, hmtf=True, L=128., K=8,
                spread='elfouhaily')
    assert sc.use_hmtf
    ds = DeviceScene(sc)
    assert 'hMTF' in ds.fields_needed()
    p0, p1 = 400, 470
    raw, nrcs = ds.run(p0, p1, chunk_snaps=3)
    torch.cuda.synchronize()
    got = raw[:, :, p0:p1].cpu().numpy().astype(np.complex128)
    ref, ref_nrcs = osar.simulate(sc, p0, p1)
    e, ph = rel_rms(got, ref), phase_rms(got, ref)
    print('hMTF %s rel RMS %.3e phase RMS %.3e' % (pol, e, ph))
    assert e < 1e-4 and ph < 1e-3
    print('NRCS max rel diff %.2e' % np.max(np.abs(nrcs[:, p0:p1] / ref_nrcs - 1)))
    np.testing.assert_allclose(nrcs[:, p0:p1], ref_nrcs, rtol=1e-6)
    sc.use_hmtf = False
    _, plain = DeviceScene(sc).run(p0, p1, chunk_snaps=3)
    assert np.max(np.abs(plain[:, p0:p1] / nrcs[:, p0:p1] - 1)) > 1e-3


def test_overlapped_pipeline_matches_serial(need_gpu, tmp_path_factory):
    """Two-stream pipeline (synthesis/Bragg of chunk c+1 beside the accumulation of
    chunk c, double-buffered) gives the single-stream result, and the oracle's."""
    import torch
    import oracle.sar as osar
    from oceansar_amd.engine import DeviceScene
    sc = _scene(tmp_path_factory, 'ov', num_ch=2, b_ati=[0., 10.])
    ds = DeviceScene(sc)
    serial, n_s = ds.run(0, 600, chunk_snaps=3)
    serial = serial.cpu().numpy()
    for rep in range(2):
        over, n_o = ds.run(0, 600, chunk_snaps=3, overlap=True)
        torch.cuda.synchronize()
        assert rel_rms(over.cpu().numpy(), serial) < 1e-6, rep
        assert np.allclose(n_o, n_s, rtol=1e-12)
    p0, p1 = 470, 590
    raw, _ = ds.run(p0, p1, chunk_snaps=2, overlap=True)
    got = raw[:, :, p0:p1].cpu().numpy().astype(np.complex128)
    ref, _ = osar.simulate(sc, p0, p1)
    assert rel_rms(got, ref) < 1e-4 and phase_rms(got, ref) < 1e-3
