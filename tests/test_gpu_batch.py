"""GPU: the batching and serving forms give every scene the plain path's result.

Batching: sequential, concurrent-stream and captured-graph runs of a small
mixed workload agree scene by scene (float32 rounding of the atomic
accumulation order), every graph replay included, and a batched scene
matches the FP64 oracle. Serving: each job's raw cube is the same alone and
in a mix of jobs arriving while others run.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _mixed_scenes(tmp_path_factory, n, seed, sides=(64, 128), pulses=(96, 260), max_ch=3):
    from oceansar_amd import batch, configs
    kws = batch.mixed_workload(n=n, seed=seed, sides=sides, pulses=pulses, max_ch=max_ch, K=8)
    d = tmp_path_factory.mktemp('mix%d' % seed)
    return [configs.load_scene(configs.cfg_text(**kw), d, name='s%d.cfg' % i)
            for i, kw in enumerate(kws)]


def _host(results):
    return [raw.cpu().numpy() for raw, _ in results]


def test_batch_modes_agree(need_gpu, tmp_path_factory):
    import torch
    from oceansar_amd.batch import Autotuner, BatchRunner
    scenes = _mixed_scenes(tmp_path_factory, 5, seed=11)
    tuner = Autotuner(str(tmp_path_factory.mktemp('tune') / 'tune.json'))
    br = BatchRunner(scenes, n_streams=3, tuner=tuner)
    assert len(tuner.cache) >= 1
    br.run_sequential()
    ref = _host(br.results())
    nrcs_ref = [n for _, n in br.results()]
    br.run_concurrent()
    for i, got in enumerate(_host(br.results())):
        assert rel_rms(got, ref[i]) < 1e-6, ('concurrent', i)
    br.capture()
    for rep in range(3):
        br.replay()
        got = _host(br.results())
        for i in range(len(scenes)):
            assert rel_rms(got[i], ref[i]) < 1e-6, ('graph replay', rep, i)
            assert np.allclose(br.results()[i][1], nrcs_ref[i], rtol=1e-9)
        # eager work between replays
        br.run_concurrent()
        torch.cuda.synchronize()


@pytest.mark.parametrize('workgroups', [0, 8])
def test_persistent_batch_agrees_with_plain(need_gpu, tmp_path_factory, workgroups):
    """osar_echo_accumulate_batch: one persistent accumulator launch per round
    over the chunks of every scene (1-3 channels, uneven sizes, several
    rounds; a grid of 8 workgroups makes each drain long queues), eager and
    as graph replays with eager work between them."""
    import torch
    from oceansar_amd.batch import BatchRunner, PersistentBatch
    scenes = _mixed_scenes(tmp_path_factory, 5, seed=13)
    br = BatchRunner(scenes, n_streams=1)
    br.run_sequential()
    ref = _host(br.results())
    nrcs_ref = [n for _, n in br.results()]
    pb = PersistentBatch(scenes, n_streams=3, chunk=4, workgroups=workgroups)
    pb.run()
    assert pb.rounds > 1
    for i, got in enumerate(_host(pb.results())):
        assert rel_rms(got, ref[i]) < 1e-6, ('persistent', i)
        np.testing.assert_allclose(pb.results()[i][1], nrcs_ref[i], rtol=1e-9)
    pb.capture()
    for rep in range(3):
        pb.replay()
        got = _host(pb.results())
        for i in range(len(scenes)):
            assert rel_rms(got[i], ref[i]) < 1e-6, ('persistent graph replay', rep, i)
        br.run_sequential()
        torch.cuda.synchronize()


def test_batched_scene_matches_oracle(need_gpu, tmp_path_factory):
    import oracle.c_oracle as co
    import oracle.sar as osar
    from oceansar_amd.batch import BatchRunner
    scenes = _mixed_scenes(tmp_path_factory, 3, seed=12)
    br = BatchRunner(scenes, n_streams=3)
    br.run_concurrent()
    co.build()
    for i, (raw, _) in enumerate(br.results()):
        sc = scenes[i]
        got = raw.cpu().numpy().astype(np.complex128)
        ref, _ = co.simulate_c(sc, 0, sc.az_steps, snaps=osar.SnapshotCache(sc))
        e = rel_rms(got, ref)
        print('batched scene %d (%dx%d, %d ch, %d pulses) rel RMS %.3e'
              % (i, sc.Nx, sc.Ny, sc.num_ch, sc.az_steps, e))
        assert e < 1e-4


def test_serve_alone_and_in_mix(need_gpu, tmp_path_factory):
    from oceansar_amd.serve import Job, Server, poisson_arrivals
    scenes = _mixed_scenes(tmp_path_factory, 6, seed=13)
    arrivals = poisson_arrivals(len(scenes), rate=200., seed=13)
    jobs = [Job(i, sc, a) for i, (sc, a) in enumerate(zip(scenes, arrivals))]
    srv = Server(n_slots=3, hbm_budget=8e9, chunk=8)
    st = srv.serve(jobs)
    assert st['jobs'] == len(jobs)
    assert all(j.t_done is not None and j.t_done >= j.arrival for j in jobs)
    assert st['p99_latency_s'] >= st['p50_latency_s'] > 0
    # the same job alone on a fresh server
    for k in (0, 3):
        alone = Job(99, scenes[k], 0.)
        Server(n_slots=1, hbm_budget=8e9, chunk=8).serve([alone])
        assert rel_rms(jobs[k].raw, alone.raw) < 1e-6, k
        assert np.allclose(jobs[k].nrcs, alone.nrcs, rtol=1e-9)


def test_serve_admission_respects_budget(need_gpu, tmp_path_factory):
    """With a budget that fits one job at a time, jobs run one after another."""
    from oceansar_amd.batch import footprint_bytes, workspace_bytes
    from oceansar_amd.serve import Job, Server
    scenes = _mixed_scenes(tmp_path_factory, 3, seed=14, sides=(128,), pulses=(96, 128))
    one = max(footprint_bytes(sc, 8) + workspace_bytes(sc, 8) for sc in scenes)
    jobs = [Job(i, sc, 0.) for i, sc in enumerate(scenes)]
    Server(n_slots=3, hbm_budget=1.2 * one, chunk=8).serve(jobs)
    spans = sorted((j.t_admit, j.t_done) for j in jobs)
    for (a0, d0), (a1, d1) in zip(spans, spans[1:]):
        assert a1 >= d0 - 1e-3     # admitted only after the previous one finished
