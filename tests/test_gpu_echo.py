"""GPU parity: Bragg facet kernel, echo accumulator and chan_profile deposit vs the FP64 oracle.

Inputs are the oracle's FP64-synthesised surface snapshots (stored float32 as
OceanSurface stores them), so these tests isolate the facet and echo kernels.
Tolerances: Bragg amplitudes come from a float32 evaluation of the facet
model (rel 1e-5 of the FP64 oracle's); the raw cube must be within 1e-4 relative RMS of the FP64 oracle
(the parity bound of the path), and is expected near 1e-6.
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


@pytest.fixture(scope='module')
def cfg1_scene(tmp_path_factory):
    from oceansar_amd import configs
    return configs.load_scene(configs.workload('small'), tmp_path_factory.mktemp('cfg1'))


def _snapshots(sc, m0, m1):
    """Oracle snapshots m0..m1 inclusive, stacked per field as device tensors."""
    import torch
    import oracle.surface as osurf
    snaps = [osurf.synth(sc.surface, sc.snapshot_time(m), ('D', 'Diff')) for m in range(m0, m1 + 1)]
    out = {}
    for n in ('Dx', 'Dy', 'Dz', 'Diffx', 'Diffy'):
        # device fields are x-major: [snapshot][x][y]
        out[n] = torch.as_tensor(np.ascontiguousarray(np.stack([s[n].T for s in snaps])),
                                 device='cuda')
    return snaps, out


def _bragg_compare(sc, m):
    import torch
    import oracle.sar as osar
    from oceansar_amd.engine import DeviceScene
    ds = DeviceScene(sc)
    snaps, dev = _snapshots(sc, m, m)
    p = m * sc.K
    npol = len(sc.pols)
    amp = torch.empty((npol, 2, sc.Nx, sc.Ny), dtype=torch.float32, device='cuda')
    tot = torch.zeros(npol, dtype=torch.float64, device='cuda')
    ds.bragg(dev['Diffx'][0], dev['Diffy'][0], p, amp, tot)
    cache = osar.SnapshotCache(sc)
    cache.cache[m] = snaps[0]
    rcs_all = osar.bragg_rcs(sc, cache, p)
    for ip, pol in enumerate(sc.pols):
        rcs = rcs_all[pol]
        ref = np.sqrt(rcs)
        got = amp[ip].cpu().numpy().transpose(0, 2, 1)
        assert np.all(np.isfinite(got))
        err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-30)
        # the facet model runs in float32 from the float32 slopes: sqrt(RCS)
        # within 1e-5 of the FP64 oracle's; the RCS sums are FP64
        print('Bragg %s: max rel err of sqrt(RCS) %.2e, RCS sum rel err %.2e'
              % (pol, np.max(err[ref > 1e-12]), abs(tot[ip].item() - rcs.sum()) / rcs.sum()))
        assert np.max(err[ref > 1e-12]) < 1e-5, pol
        assert np.all(got[ref <= 1e-12] < 1e-5), pol
        assert abs(tot[ip].item() - rcs.sum()) / rcs.sum() < 1e-6, pol


def test_bragg_kernel_matches_oracle(need_gpu, cfg1_scene):
    """cos^2 spreading, vv."""
    _bragg_compare(cfg1_scene, 25)


@pytest.mark.parametrize('pol', ['vv', 'hh', 'DP'])
def test_bragg_kernel_elfouhaily(need_gpu, tmp_path_factory, pol):
    """Elfouhaily spreading (the Delta(k) terms), every polarisation including the
    DP mask quirk, wind across the scene so both half-planes are populated."""
    from oceansar_amd import configs
    sc = configs.load_scene(configs.workload('small', L=128., spread='elfouhaily', pol=pol,
                                             wind_dir=60., num_ch=1),
                            tmp_path_factory.mktemp('be' + pol))
    _bragg_compare(sc, 7)


def test_accumulator_matches_oracle(need_gpu, cfg1_scene):
    import torch
    import oracle.sar as osar
    from oceansar_amd.engine import DeviceScene
    sc = cfg1_scene
    p0, n = 392, 40                      # crosses two snapshot boundaries (K = 16)
    m0, m1 = p0 // sc.K, (p0 + n - 1) // sc.K + 1
    snaps, dev = _snapshots(sc, m0, m1)
    ds = DeviceScene(sc)
    r0, r1 = int(sc.rcs_slot[p0]), int(sc.rcs_slot[p0 + n - 1])
    amp = torch.empty((r1 - r0 + 1, 1, 2, sc.Nx, sc.Ny), dtype=torch.float32, device='cuda')
    for r in range(r0, r1 + 1):
        pr = int(sc.rcs_refresh[r])
        ds.bragg(dev['Diffx'][pr // sc.K - m0], dev['Diffy'][pr // sc.K - m0], pr, amp[r - r0])
    raw = ds.new_raw()
    ds.accumulate(dev['Dx'], dev['Dy'], dev['Dz'], m0, amp, r0, p0, n, raw)
    torch.cuda.synchronize()
    got = raw[:, :, p0:p0 + n].cpu().numpy().astype(np.complex128)
    cache = osar.SnapshotCache(sc)
    for i, m in enumerate(range(m0, m1 + 1)):
        cache.cache[m] = snaps[i]
    ref, _ = osar.simulate(sc, p0, p0 + n, snaps=cache)
    e = rel_rms(got, ref)
    print('accumulator rel RMS %.3e phase RMS %.3e rad' % (e, phase_rms(got, ref)))
    assert e < 1e-4
    assert phase_rms(got, ref) < 1e-3
    # pulses outside the window stay zero
    assert float(raw[:, :, :p0].abs().max()) == 0.0


def test_accumulator_launch_shapes_agree(need_gpu, cfg1_scene):
    """Different (P, R, W, columns) decompositions give the same lines (small windows
    force the global-memory fallback path)."""
    import torch
    from oceansar_amd.engine import DeviceScene, LaunchShape
    sc = cfg1_scene
    p0, n = 500, 24
    m0, m1 = p0 // sc.K, (p0 + n - 1) // sc.K + 1
    _, dev = _snapshots(sc, m0, m1)
    ds = DeviceScene(sc)
    r0, r1 = int(sc.rcs_slot[p0]), int(sc.rcs_slot[p0 + n - 1])
    amp = torch.empty((r1 - r0 + 1, 1, 2, sc.Nx, sc.Ny), dtype=torch.float32, device='cuda')
    for r in range(r0, r1 + 1):
        pr = int(sc.rcs_refresh[r])
        ds.bragg(dev['Diffx'][pr // sc.K - m0], dev['Diffy'][pr // sc.K - m0], pr, amp[r - r0])
    outs = []
    # column chunks: the window follows the chunk's first column (sr_col is monotonic)
    shapes = (ds.shape, LaunchShape(2, 7, 64), LaunchShape(4, 256, 32), LaunchShape(2, 16, 128),
              LaunchShape(8, 100, 96), LaunchShape(4, 128, ds.window_bins(128, 64), 64),
              LaunchShape(2, 32, ds.window_bins(32, 100), 100), LaunchShape(4, 64, 24, 50),
              LaunchShape(16, 32, ds.window_bins(32, 64), 64), LaunchShape(16, 40, 40, 0))
    for shape in shapes:
        raw = ds.new_raw()
        ds.accumulate(dev['Dx'], dev['Dy'], dev['Dz'], m0, amp, r0, p0, n, raw, shape=shape)
        outs.append(raw[:, :, p0:p0 + n].cpu().numpy().astype(np.complex128))
    for o in outs[1:]:
        assert rel_rms(o, outs[0]) < 2e-6


def test_accumulator_without_scattering_adds_nothing(need_gpu, cfg1_scene):
    """Neither Bragg nor specular scattering: the call adds nothing (the plain
    accumulator, which assumes the Bragg flag, is not launched) and the lines
    keep what they held."""
    import torch
    from oceansar_amd.engine import DeviceScene
    sc = cfg1_scene
    p0, n = 500, 24
    m0, m1 = p0 // sc.K, (p0 + n - 1) // sc.K + 1
    _, dev = _snapshots(sc, m0, m1)
    ds = DeviceScene(sc)
    raw = ds.new_raw()
    raw.real.fill_(1.5)
    raw.imag.fill_(-0.25)
    before = raw.clone()
    saved = sc.bragg_enable
    try:
        sc.bragg_enable = False
        ds.accumulate(dev['Dx'], dev['Dy'], dev['Dz'], m0, None, 0, p0, n, raw)
    finally:
        sc.bragg_enable = saved
    torch.cuda.synchronize()
    assert torch.equal(raw, before)


def test_accumulator_rejects_scatterer_series_beyond_32bit_offsets(need_gpu, cfg1_scene):
    """The scatterer gathers use 32-bit byte offsets: a series of more than 2^29
    entries, or one too short for the pulses, is refused before any launch."""
    import ctypes
    import torch
    from oceansar_amd import _lib
    from oceansar_amd.engine import DeviceScene
    sc = cfg1_scene
    ds = DeviceScene(sc)
    p0, n = 0, 8
    m1 = (p0 + n - 1) // sc.K + 1
    f = torch.zeros((m1 + 1, sc.Nx, sc.Ny), dtype=torch.float32, device='cuda')
    amp = torch.zeros((1, 1, 2, sc.Nx, sc.Ny), dtype=torch.float32, device='cuda')
    raw = ds.new_raw()
    d = ds.echo_desc(f, f, f, 0, amp, int(sc.rcs_slot[p0]), p0, n, raw)
    assert d.scat_len == ds.scat_m.numel() > ds.n_pulses
    lib = ds.ctx.lib
    # the descriptor is otherwise valid (zero amplitudes: nothing is added)
    assert lib.osar_echo_accumulate(ds.ctx.h, ctypes.byref(d), _lib.stream_ptr()) == 0
    for bad in ((1 << 29) + 1, ds.n_pulses):
        d.scat_len = bad
        assert lib.osar_echo_accumulate(ds.ctx.h, ctypes.byref(d), _lib.stream_ptr()) != 0
        assert 'scat_len' in lib.osar_last_error().decode()
    torch.cuda.synchronize()
    assert float(raw.abs().max()) == 0.0


def test_deposit_matches_oracle(need_gpu):
    import torch
    import oracle.echo as oecho
    from oceansar_amd.range_profile import calc_sinc_vec, chan_profile
    rng = np.random.default_rng(7)
    n, nbins = 20000, 300
    srg = rng.uniform(-20., nbins * 1.3 + 20., n)
    scene = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    tab = calc_sinc_vec(8, 20, Fs=1.1)
    ref = np.zeros(nbins, dtype=np.complex128)
    oecho.chan_profile(srg, scene, 1.3, -2.73, tab, 8, 20, ref)
    out = torch.zeros(nbins, dtype=torch.complex64, device='cuda')
    chan_profile(torch.as_tensor(srg, device='cuda'),
                 torch.as_tensor(scene.astype(np.complex64), device='cuda'),
                 1.3, -2.73, torch.as_tensor(tab, device='cuda'), 8, 20, out)
    got = out.cpu().numpy()
    assert rel_rms(got, ref) < 1e-5
