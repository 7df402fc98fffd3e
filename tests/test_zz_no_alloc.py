"""osar_surface_synth_noalloc: the synthesis with its workspace in the library's
per-device scratch array instead of device memory it allocates.

Through the C ABI: two calls in a row (the second on a smaller surface with
fewer snapshots and another field set, so the scratch holds the previous call's
contents) give the same bits as osar_surface_synth on the same descriptors, and
a descriptor whose workspace exceeds the scratch is rejected before any launch.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _desc(ds, sc, times_t, out, names, slope_fp32):
    from oceansar_amd import _lib
    from oceansar_amd.engine import mirror_grid
    d = _lib.SurfaceDesc()
    d.nx, d.ny = sc.Nx, sc.Ny
    d.fields = sum(_lib.FIELD_BITS[n] for n in names)
    d.choppy = 0
    d.n_times = times_t.numel()
    d.mirror_grid = 1 if (mirror_grid(sc.surface.kx_s) and mirror_grid(sc.surface.ky_s)) else 0
    d.slope_fp32 = 1 if slope_fp32 else 0
    d.depth = -1.0 if sc.surface.depth is None else float(sc.surface.depth)
    d.current_x, d.current_y = float(sc.surface.current[0]), float(sc.surface.current[1])
    d.kx_res, d.ky_res = float(sc.surface.kx_res), float(sc.surface.ky_res)
    d.kx, d.ky, d.coefs = _lib.ptr(ds.kx), _lib.ptr(ds.ky), _lib.ptr(ds.coefs)
    d.times = _lib.ptr(times_t)
    for i, n in enumerate(_lib.FIELD_ORDER):
        if n in names:
            d.out[i] = out[n].data_ptr()
    return d


def _run(ds, sc, times, names, slope_fp32, noalloc):
    import torch
    from oceansar_amd import _lib
    t = torch.as_tensor(np.asarray(times, np.float64), device='cuda')
    out = {n: torch.full((len(times), sc.Nx, sc.Ny), np.nan, dtype=torch.float32, device='cuda')
           for n in names}
    d = _desc(ds, sc, t, out, names, slope_fp32)
    fn = ds.ctx.lib.osar_surface_synth_noalloc if noalloc else ds.ctx.lib.osar_surface_synth
    _lib.check(fn(ds.ctx.h, ctypes.byref(d), _lib.stream_ptr()))
    torch.cuda.synchronize()
    return {n: out[n].cpu().numpy() for n in names}, d


def test_noalloc_matches_synth_two_calls(need_gpu, tmp_path_factory):
    import torch
    from oceansar_amd import _lib, configs
    from oceansar_amd.engine import DeviceScene
    big = configs.load_scene(configs.workload('small', L=512., dx=1.), tmp_path_factory.mktemp('n1'))
    small = configs.load_scene(configs.workload('small', L=128., dx=1., seed=7),
                               tmp_path_factory.mktemp('n2'))
    cases = [(big, [0.0, 1.5, 7.25], ['Dx', 'Dy', 'Dz', 'Diffx', 'Diffy'], True),
             (small, [2.0, 3.0], ['Dz', 'Diffx', 'Diffy', 'Diffxx', 'Diffyy', 'Diffxy'], False)]
    for sc, times, names, f32 in cases:
        ds = DeviceScene(sc)
        got, d = _run(ds, sc, times, names, f32, noalloc=True)
        need = ds.ctx.lib.osar_surface_workspace_bytes(ctypes.byref(d))
        assert 0 < need <= _lib.SYNTH_SCRATCH_BYTES
        ref, _ = _run(ds, sc, times, names, f32, noalloc=False)
        for n in names:
            assert np.isfinite(got[n]).all(), (sc.Nx, n)
            assert np.array_equal(got[n], ref[n]), (sc.Nx, n)
        print('noalloc %dx%d, %d snapshots, %d fields (%d B of scratch): bit-equal'
              % (sc.Nx, sc.Ny, len(times), len(names), need))
    torch.cuda.synchronize()


def test_noalloc_rejects_workspace_beyond_scratch(need_gpu, tmp_path_factory):
    import torch
    from oceansar_amd import _lib, configs
    from oceansar_amd.engine import DeviceScene
    sc = configs.load_scene(configs.workload('small', L=1024., dx=1.), tmp_path_factory.mktemp('n3'))
    ds = DeviceScene(sc)
    names = ['Dx', 'Dy', 'Dz', 'Diffx', 'Diffy']
    n = 40                                   # 40 snapshots of 1024^2: 1.7 GB of workspace
    t = torch.zeros(n, dtype=torch.float64, device='cuda')
    out = {k: torch.zeros((n, sc.Nx, sc.Ny), dtype=torch.float32, device='cuda') for k in names}
    d = _desc(ds, sc, t, out, names, True)
    assert ds.ctx.lib.osar_surface_workspace_bytes(ctypes.byref(d)) > _lib.SYNTH_SCRATCH_BYTES
    rc = ds.ctx.lib.osar_surface_synth_noalloc(ds.ctx.h, ctypes.byref(d), _lib.stream_ptr())
    assert rc == -1                          # OSAR_EINVAL
    assert 'scratch' in ds.ctx.lib.osar_last_error().decode()
    torch.cuda.synchronize()
    assert float(out['Dz'].abs().max()) == 0.0      # nothing was launched
