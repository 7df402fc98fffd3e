"""Host logic: cfg parsing, scene sizing, cadence schedule, C-ABI exports, C oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cfg_parser_reference_semantics(tmp_path):
    from oceansar_amd.cfg import ConfigFile
    text = """# comment
; another
[sar]
f0=5.4e9   ; inline comment
num_ch=2
pol=DP
flag=True
b_ati=[0, 10.5]
names=['a','b']
bools=[True, False]
big=1e+3
v_ground=auto
[ocean]
Lx=512
"""
    p = tmp_path / 'x.cfg'
    p.write_text(text)
    c = ConfigFile(str(p))
    assert c.sar.f0 == 5.4e9 and c.sar.num_ch == 2 and isinstance(c.sar.num_ch, int)
    assert c.sar.pol == 'DP' and c.sar.flag is True and c.sar.v_ground == 'auto'
    np.testing.assert_array_equal(c.sar.b_ati, [0., 10.5])
    assert c.sar.names == ['a', 'b']
    np.testing.assert_array_equal(c.sar.bools, [True, False])
    assert c.sar.big == '1e'           # '+' is outside the value class (cfg.py:100)
    assert c.ocean.Lx == 512
    assert c.sections == [c, 'sar', 'ocean']
    c.save(str(tmp_path / 'y.cfg'))
    c2 = ConfigFile(str(tmp_path / 'y.cfg'))
    assert c2.sar.num_ch == 2 and c2.sar.pol == 'DP'


@pytest.mark.parametrize('idx,nx,pulses,orig,samp,nscat', [
    ('small', 256, 1024, 115, 128, 38400), ('ati_swell', 1024, 4096, 267, 512, 54000),
    ('large', 2048, 8192, 624, 1024, 54000)])
def test_workload_sizes(tmp_path, idx, nx, pulses, orig, samp, nscat):
    """Sizes derived by the reference formulas (sar_raw_nompi.py:264-278)."""
    from oceansar_amd import configs
    sc = configs.load_scene(configs.workload(idx), tmp_path)
    assert (sc.Nx, sc.Ny, sc.az_steps) == (nx, nx, pulses)
    # the pulse count also follows from t_span without the n_pulses override
    assert int(np.floor(sc.t_span / sc.t_step)) == pulses
    assert (sc.rg_samp_orig, sc.rg_samp) == (orig, samp)
    assert sc.scat_m.N == nscat and sc.scat_p.N == nscat
    assert sc.min_sr == -2.73
    assert 0 <= sc.scat_m.ind0.min() and sc.scat_m.ind0.max() < sc.scat_m.N


def test_rcs_schedule_float_cadence(tmp_path):
    """Without rcs_update_every the refresh follows t_now - t_last > dt (sar_raw_nompi.py:432)."""
    from oceansar_amd import configs
    text = configs.workload('small', K=1, dt=0.05)
    sc = configs.load_scene(text, tmp_path)
    ref, t_last = [], -1.
    for p in range(sc.az_steps):
        t = p * sc.t_step
        if t - t_last > 0.05:
            ref.append(p)
            t_last = t
    np.testing.assert_array_equal(sc.rcs_refresh, ref)
    assert sc.rcs_slot[ref[3]] == 3 and sc.rcs_slot[ref[3] - 1] == 2


def test_seeded_inputs_reproducible(tmp_path):
    from oceansar_amd import configs
    a = configs.load_scene(configs.workload('small'), tmp_path)
    b = configs.load_scene(configs.workload('small'), tmp_path)
    c = configs.load_scene(configs.workload('small', seed=9), tmp_path)
    np.testing.assert_array_equal(a.surface.wave_coefs, b.surface.wave_coefs)
    np.testing.assert_array_equal(a.scat_m.ind0, b.scat_m.ind0)
    assert not np.array_equal(a.surface.wave_coefs, c.surface.wave_coefs)


def test_surface_statistics(tmp_path):
    """E[var Dz] = sum of the directional spectrum; slopes' variance ~ rmss^2 (ocean.py:181)."""
    from oceansar_amd import configs
    import oracle.surface as osurf
    sc = configs.load_scene(configs.workload('small'), tmp_path)
    s = sc.surface
    f = osurf.synth(s, 0.0, ('D', 'Diff'))
    dk2 = float(s.kx_res * s.ky_res)
    var_z = np.sum(s.wave_dirspec) * dk2
    # few long-wave modes dominate the elevation variance in a 256 m patch: loose bound
    assert abs(np.var(f['Dz']) / var_z - 1) < 0.5
    assert abs(np.sqrt(np.mean(f['Diffx'].astype(float) ** 2)) / s.rmss_x - 1) < 0.1


def _header_functions():
    txt = open(os.path.join(ROOT, 'include', 'oceansar_hip.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(osar_[a-z0-9_]+)\s*\(', txt)))


def test_library_exports_every_declared_symbol():
    from oceansar_amd import _lib
    lib = _lib.load()                      # loads on a CPU-only host; no GPU call made
    names = _header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTS)
    assert lib.osar_version() >= 100
    for i, cls in enumerate((_lib.EchoDesc, _lib.BraggDesc, _lib.SurfaceDesc, _lib.FocusDesc,
                             _lib.InitDesc, _lib.CoregDesc, _lib.MultilookDesc)):
        assert lib.osar_struct_size(i) == ctypes.sizeof(cls)


def test_product_path_fails_loudly_without_gpu():
    import torch
    from oceansar_amd import _lib
    if torch.cuda.is_available():
        pytest.skip('GPU visible')
    with pytest.raises(RuntimeError, match='GPU'):
        _lib.Context(0)


def test_c_oracle_matches_numpy_oracle(tmp_path):
    from oceansar_amd import configs
    import oracle.c_oracle as co
    import oracle.sar as osar
    sc = configs.load_scene(configs.workload('small', L=96., pol='DP', num_ch=2, b_ati=[0, 10],
                                             b_xti=[0, 25.], K=4), tmp_path)
    cache = osar.SnapshotCache(sc)
    a, na = osar.simulate(sc, 300, 330, snaps=cache)
    b, nb = co.simulate_c(sc, 300, 330, snaps=cache)
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-12
    np.testing.assert_allclose(na, nb, rtol=1e-12)


def test_arange_grid_detection():
    """The Bragg table search computes grid points only for exact numpy.arange grids."""
    from oceansar_amd.engine import arange_grid
    k = np.arange(1, 2 * 113.27, 0.1)
    k0, dk, uni = arange_grid(k)
    assert uni == 1 and k0 == 1.0 and dk == k[1] - k[0]
    assert np.array_equal(k0 + np.arange(k.size) * dk, k)
    bent = k.copy()
    bent[17] = np.nextafter(bent[17], 0.)
    assert arange_grid(bent)[2] == 0
    assert arange_grid(np.linspace(1, 5, 9) ** 2)[2] == 0
    assert arange_grid(np.array([3.]))[2] == 0


def test_frozen_ocean_holds_the_surface_at_t0(tmp_path):
    """[ocean] frozen_ocean: every snapshot is the surface at t = 0 (the reference
    skips the time update, sar_raw_nompi.py:357-358); the Bragg phase still
    advances with the pulse index."""
    from oceansar_amd import configs
    import oracle.sar as osar
    sc = configs.load_scene(configs.workload('small', L=64., n_pulses=128, K=4, frozen=True),
                            tmp_path)
    assert sc.frozen and all(sc.snapshot_time(m) == 0. for m in range(sc.n_snapshots))
    cache = osar.SnapshotCache(sc)
    a, b = cache.get(0), cache.get(7)
    for n in a:
        np.testing.assert_array_equal(a[n], b[n])
    live = configs.load_scene(configs.workload('small', L=64., n_pulses=128, K=4), tmp_path)
    assert not live.frozen and live.snapshot_time(7) == 28 * live.t_step
    r0, _ = osar.simulate(sc, 40, 44, snaps=cache)
    assert np.all(np.isfinite(r0)) and np.abs(r0).max() > 0


def test_workspace_bytes_match_the_library(tmp_path):
    """batch.workspace_bytes (admission and launch-group sizing) agrees with
    osar_surface_workspace_bytes for the fields each kind of scene synthesises."""
    import ctypes
    from oceansar_amd import _lib, batch, configs
    from oceansar_amd.engine import DeviceScene
    from oceansar_amd.engine import spectra_bytes
    lib = _lib.load()
    for kw in (dict(), dict(specular=True, base='ati_swell', L=64.), dict(hmtf=True),
               dict(specular=True, choppy=True, base='ati_swell', L=64.),
               dict(choppy=True, hmtf=True)):
        base = kw.pop('base', 'small')
        sc = configs.load_scene(configs.workload(base, L=kw.pop('L', 128.), **kw), tmp_path)
        names = DeviceScene.fields_needed(types_ns(sc))
        for chunk in (1, 8, 32):
            d = _lib.SurfaceDesc()
            d.nx, d.ny, d.n_times = sc.Nx, sc.Ny, chunk + 1
            d.fields = sum(_lib.FIELD_BITS[n] for n in names)
            d.choppy = 1 if sc.surface.choppy else 0
            d.slope_fp32 = 1 if (sc.bragg_enable and not sc.spec_enable) else 0
            assert lib.osar_surface_workspace_bytes(ctypes.byref(d)) == \
                batch.workspace_bytes(sc, chunk), (kw, chunk)
    # every field combination's plan, against the library's
    rng = np.random.default_rng(0)
    for _ in range(300):
        names = [n for n in _lib.FIELD_ORDER if rng.random() < 0.4]
        if not names:
            continue
        choppy, f32 = bool(rng.random() < 0.5), bool(rng.random() < 0.5)
        d = _lib.SurfaceDesc()
        d.nx, d.ny, d.n_times = 64, 32, 1
        d.fields = sum(_lib.FIELD_BITS[n] for n in names)
        d.choppy, d.slope_fp32 = int(choppy), int(f32)
        want = lib.osar_surface_workspace_bytes(ctypes.byref(d))
        # the spectra, then the dispersion and 1/|k| tables
        assert want == spectra_bytes(names, choppy, f32) * 64 * 32 + 16 * 64 * 32, \
            (names, choppy, f32)


def types_ns(sc):
    import types
    return types.SimpleNamespace(sc=sc)
