"""GPU parity at the full sizes of the two-channel ATI workload and of the
batching and serving mixes.

ATI + swell (1024 x 1024 facets at 0.5 m, 35 deg, 2 channels 0/10 m, 200 m
swell, Bragg + Kirchhoff specular, K = 8, 4096 pulses): every synthesised
field at 512^2 and 1024^2 against numpy's complex128 ifft2, and a 12-pulse
window across a snapshot and a Bragg refresh against the FP64 numpy oracle
(relative RMS <= 1e-4, phase RMS <= 1 mrad, NRCS).

Batching mix (64 scenes of sides 128-1024, 512-4096 pulses, 1-4 channels,
seed 4): sequential, concurrent and captured-graph runs agree scene by scene,
and the largest 1024^2 scene's window matches the C oracle.
Serving mix (seed 5, 10 % of jobs at 2048^2 with 2 channels, 200 GB budget):
each job's result is the same alone and in the mix.
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
def ati_scene(tmp_path_factory):
    from oceansar_amd import configs
    sc = configs.load_scene(configs.workload('ati_swell'), tmp_path_factory.mktemp('ati'))
    assert (sc.Nx, sc.Ny, sc.num_ch, sc.az_steps, sc.K) == (1024, 1024, 2, 4096, 8)
    assert sc.spec_enable and sc.bragg_enable
    return sc


NAMES = ['Dx', 'Dy', 'Dz', 'Diffx', 'Diffy', 'Diffxx', 'Diffyy', 'Diffxy']


def _synth_compare(sc, times):
    import torch
    import oracle.surface as osurf
    from oceansar_amd.engine import DeviceScene
    ds = DeviceScene(sc)
    out = {n: torch.empty((len(times), sc.Nx, sc.Ny), dtype=torch.float32, device='cuda')
           for n in NAMES}
    ds.synth(times, out, NAMES)
    torch.cuda.synchronize()
    for ti, t in enumerate(times):
        ref = osurf.synth(sc.surface, t, ('D', 'Diff', 'Diff2'))
        for n in NAMES:
            got = out[n][ti].cpu().numpy().T
            r = ref[n]
            same = np.mean(got == r)
            err = np.max(np.abs(got - r)) / np.sqrt(np.mean(r.astype(np.float64) ** 2))
            print('%dx%d t=%g %s: %.5f bit-equal, max err %.2e RMS' % (sc.Nx, sc.Ny, t, n, same, err))
            assert same > 0.999, (n, t, same)
            assert err < 1e-6, (n, t, err)


def test_synth_1024_all_fields(need_gpu, ati_scene):
    """1024-point rows and columns (radix 8, 8, 4, 4; column blocks of 4 in runs of 8)."""
    _synth_compare(ati_scene, [0.0, 1137.6 / 1800.])


def test_synth_512_all_fields(need_gpu, tmp_path_factory):
    from oceansar_amd import configs
    sc = configs.load_scene(configs.workload('ati_swell', L=256.),
                            tmp_path_factory.mktemp('a512'))
    assert (sc.Nx, sc.Ny) == (512, 512)
    _synth_compare(sc, [0.0, 2.71])


def test_ati_swell_window_matches_oracle(need_gpu, ati_scene):
    import torch
    import oracle.sar as osar
    from oceansar_amd.engine import DeviceScene
    sc = ati_scene
    ds = DeviceScene(sc)
    p0, p1 = 2044, 2056                     # snapshot 256 and the refresh at 2048
    assert 2048 in set(int(r) for r in sc.rcs_refresh)
    raw, nrcs = ds.run(p0, p1, chunk_snaps=2)
    torch.cuda.synchronize()
    got = raw[:, :, p0:p1].cpu().numpy().astype(np.complex128)
    snaps = osar.SnapshotCache(sc)
    ref, ref_nrcs = osar.simulate(sc, p0, p1, snaps=snaps)
    e, ph = rel_rms(got, ref), phase_rms(got, ref)
    # stationary-phase facets of the window's first pulse (sar_raw_nompi.py:401-428)
    import oracle.ka as oka
    f = osar.lerp_fields(snaps, sc, p0, ('Diffx', 'Diffy', 'Diffxx', 'Diffyy', 'Diffxy'))
    _, _, az_proj = osar.row_geometry(sc, p0)
    E = oka.field_spa(sc.k0, sc.pol, sc.inc, az_proj[:, 0], f['Diffx'], f['Diffy'],
                      f['Diffxx'], f['Diffyy'], f['Diffxy'], float(sc.surface.dx),
                      float(sc.surface.dy))
    n_sp = int(np.count_nonzero(E))
    print('ATI swell 1024^2 x 2 ch window: rel RMS %.3e phase RMS %.3e; %d stationary-phase '
          'facets of %d at pulse %d' % (e, ph, n_sp, sc.Nx * sc.Ny, p0))
    assert e < 1e-4 and ph < 1e-3
    for ch in range(sc.num_ch):
        assert rel_rms(got[:, ch], ref[:, ch]) < 1e-4, ch
    # the Bragg facet model runs in float32: its RCS sums agree to ~1e-7
    print('NRCS max rel diff %.2e' % np.max(np.abs(nrcs[:, p0:p1] / ref_nrcs - 1)))
    np.testing.assert_allclose(nrcs[:, p0:p1], ref_nrcs, rtol=1e-6)


def _mix(tmp_path_factory, kws, tag):
    from oceansar_amd import configs
    d = tmp_path_factory.mktemp(tag)
    return [configs.load_scene(configs.cfg_text(**kw), d, name='s%d.cfg' % i)
            for i, kw in enumerate(kws)]


def test_batching_mix_full_size(need_gpu, tmp_path_factory):
    import torch
    import oracle.c_oracle as co
    import oracle.sar as osar
    from oceansar_amd import batch
    scenes = _mix(tmp_path_factory, batch.mixed_workload(n=64, seed=4), 'b64')
    assert sorted(set(sc.Nx for sc in scenes)) == [128, 256, 512, 1024]
    tuner = batch.Autotuner(str(tmp_path_factory.mktemp('tune') / 'tune.json'))
    br = batch.BatchRunner(scenes, n_streams=4, tuner=tuner)
    br.run_sequential()
    ref = [(raw.cpu().numpy(), n.copy()) for raw, n in br.results()]
    br.run_concurrent()
    conc = [raw.cpu().numpy() for raw, _ in br.results()]
    br.capture()
    worst = 0.
    for rep in range(2):
        br.replay()
        torch.cuda.synchronize()
        for i, (raw, n) in enumerate(br.results()):
            e = rel_rms(raw.cpu().numpy(), ref[i][0])
            worst = max(worst, e, rel_rms(conc[i], ref[i][0]))
            assert e < 1e-6, ('graph', rep, i)
            assert np.allclose(n, ref[i][1], rtol=1e-9)
    print('64-scene mix: worst rel RMS between modes %.2e' % worst)
    # the largest scene against the C oracle on a window across a snapshot
    big = max(range(len(scenes)), key=lambda i: scenes[i].contributions)
    sc = scenes[big]
    assert sc.Nx == 1024
    p = (sc.az_steps // (2 * sc.K)) * sc.K
    p0, p1 = p - 6, p + 10
    co.build()
    want, _ = co.simulate_c(sc, p0, p1, snaps=osar.SnapshotCache(sc))
    got = ref[big][0][:, :, p0:p1].astype(np.complex128)
    e = rel_rms(got, want)
    print('scene %d (%d^2, %d ch, %d pulses) window rel RMS %.3e phase RMS %.3e'
          % (big, sc.Nx, sc.num_ch, sc.az_steps, e, phase_rms(got, want)))
    assert e < 1e-4 and phase_rms(got, want) < 1e-3


def test_serving_mix_alone_equals_in_mix(need_gpu, tmp_path_factory):
    from oceansar_amd import batch, serve
    kws = batch.mixed_workload(n=48, seed=5)
    rng = np.random.default_rng(55)
    for kw in kws:
        if rng.uniform() < 0.1:
            kw.update(L=2048., dx=1., num_ch=2, b_ati=[0., 5.])
    scenes = _mix(tmp_path_factory, kws, 's48')
    big = [i for i, sc in enumerate(scenes) if sc.Nx == 2048]
    assert big, 'the seed-5 mix holds 2048^2 jobs'
    W = sum(sc.contributions for sc in scenes)
    arrivals = serve.poisson_arrivals(len(scenes), rate=30., seed=5)
    jobs = [serve.Job(i, sc, a) for i, (sc, a) in enumerate(zip(scenes, arrivals))]
    st = serve.Server(n_slots=4, hbm_budget=200e9).serve(jobs)
    assert st['jobs'] == len(jobs)
    print('serving 48 jobs (%d at 2048^2): p50 %.3f s p99 %.3f s, %.3e contributions/s'
          % (len(big), st['p50_latency_s'], st['p99_latency_s'], W / st['makespan_s']))
    for k in big[:2] + [0, 7]:
        alone = serve.Job(99, scenes[k], 0.)
        serve.Server(n_slots=1, hbm_budget=200e9).serve([alone])
        e = rel_rms(jobs[k].raw, alone.raw)
        assert e < 1e-6, (k, e)
        assert np.allclose(jobs[k].nrcs, alone.nrcs, rtol=1e-9)
