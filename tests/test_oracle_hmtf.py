"""The hydrodynamic modulation of the Bragg RCS (use_hmtf, sar_raw_nompi.py:448-456)
in the oracle: RCS times 1 + hMTF, with hMTF below -1 set to -1 first.

Pins: the clamp and the float32 factor by hand values; a surface without waves
longer than 20 m has hMTF = 0 (ocean.py:524 cuts k > 2 pi / 20), so use_hmtf
leaves its raw data unchanged to the bit; with long waves the modulated NRCS
is the facet sum of RCS (1 + hMTF), which the plain run's RCS and the synthesised
field give independently.
"""
import numpy as np

import oracle.sar as osar
from oceansar_amd import configs


def test_modulation_clamps_at_minus_one():
    h = np.array([-3., -1., -0.5, 0., 2.5], dtype=np.float32)
    np.testing.assert_array_equal(osar.hmtf_modulation(h), [0., 0., 0.5, 1., 3.5])
    # the factor is rounded to float32 (1 + hMTF of the float32 field)
    x = np.float32(1e-8)
    assert osar.hmtf_modulation(np.array([x]))[0] == 1.0


def test_short_waves_only_leave_raw_unchanged(tmp_path):
    kw = dict(L=16., dx=1., n_pulses=64, K=4)
    a = configs.load_scene(configs.workload('small', **kw), tmp_path, 'a.cfg')
    b = configs.load_scene(configs.workload('small', hmtf=True, **kw), tmp_path, 'b.cfg')
    assert b.use_hmtf and not a.use_hmtf
    snaps = osar.SnapshotCache(b)
    assert np.all(snaps.get(3)['hMTF'] == 0)
    ra, na = osar.simulate(a, 20, 40)
    rb, nb = osar.simulate(b, 20, 40)
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(na, nb)


def test_modulated_nrcs_is_the_facet_sum(tmp_path):
    kw = dict(L=128., dx=1., n_pulses=128, K=8)
    a = configs.load_scene(configs.workload('small', **kw), tmp_path, 'a.cfg')
    b = configs.load_scene(configs.workload('small', hmtf=True, **kw), tmp_path, 'b.cfg')
    sb = osar.SnapshotCache(b)
    p = int(b.rcs_refresh[2])
    h = sb.get(p // b.K)['hMTF']
    assert 0.05 < np.std(h) < 1.0 and np.min(h) < 0
    plain = osar.bragg_rcs(a, osar.SnapshotCache(a), p)
    mod = osar.bragg_rcs(b, sb, p)
    for pol in plain:
        f = 1. + np.maximum(h.astype(np.float64), -1.)
        np.testing.assert_allclose(mod[pol], plain[pol] * f[None], rtol=1e-7, atol=0)
        assert np.sum(mod[pol]) != np.sum(plain[pol])
    _, nb = osar.simulate(b, p, p + 1)
    assert np.isclose(nb[0, 0] * b.surface_area, np.sum(mod[b.pols[0]]), rtol=1e-12)
