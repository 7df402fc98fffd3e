"""GPU parity: SAR focusing (osar_sar_focus) against the FP64 numpy oracle.

Both sides run the processing loop of sar_focus (sar_processor.py:131-225)
in complex128 on the same complex64 raw planes, so the SLCs agree to FP64
FFT rounding (relative RMS below 1e-12): a point target (the oracle's
closed-form pin), random planes at the 8192-line, 1024-sample size of the
large workload's padded raw cube, and the raw file of a GPU sar_raw run read
back and focused through sar_focus().
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _params():
    from oceansar_amd.sar_processor import FocusParams
    return FocusParams(1800., 110e6, 5.405e9, 7600., 616400.5, 1200., 1.0, 6.0, 6.0)


def _oracle(raw, fp):
    import oracle.focus as of
    return of.focus_channel(raw.astype(np.complex128), fp.prf, fp.rg_sampling, fp.f0,
                            fp.v_ground, fp.sr0, fp.doppler_bw, fp.az_weighting, fp.ant_l_tx,
                            fp.ant_l_rx)


def _gpu(raw, fp):
    import torch
    from oceansar_amd import _lib
    from oceansar_amd.sar_processor import focus_channel
    d = torch.as_tensor(raw.astype(np.complex64), device='cuda')
    return focus_channel(_lib.context(0), d, fp).cpu().numpy()


def test_point_target_focus_matches_oracle(need_gpu):
    import sys
    sys.path.insert(0, os.path.dirname(__file__))
    import test_oracle_focus as tof
    raw = (tof._target(2048, 128, 1000, 40.3) + tof._target(2048, 128, 1400, 90.0, 0.3))[None]
    fp = _params()
    raw = raw.astype(np.complex64)
    got, want = _gpu(raw, fp), _oracle(raw, fp)
    e = rel_rms(got, want)
    print('point targets 2048 x 128 -> %s: rel RMS %.2e' % (got.shape, e))
    assert got.shape == want.shape and e < 1e-12
    mag = np.abs(got[0])
    ia, ir = np.unravel_index(np.argmax(mag), mag.shape)
    assert (ia + fp.valid_rows(2048)[0], ir) == (1000, 40)


@pytest.mark.parametrize('shape', [(2, 8192, 624), (1, 1000, 115), (2, 4096, 267)])
def test_random_planes_match_oracle(need_gpu, shape):
    """Padded sizes 8192 x 1024 (the large workload's raw cube per channel),
    1024 x 128 and 4096 x 512, two polarisation planes where given."""
    rng = np.random.default_rng(sum(shape))
    raw = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    fp = _params()
    got, want = _gpu(raw, fp), _oracle(raw, fp)
    e = rel_rms(got, want)
    print('random %s: rel RMS %.2e' % (shape, e))
    assert got.shape == want.shape and e < 1e-12


def test_sar_focus_of_a_sar_raw_file(need_gpu, tmp_path):
    """sar_raw writes the raw file, sar_focus reads it as the reference does
    (sar_processor.py:71-77) and writes the SLC file; its SLC equals the
    oracle's focusing of the same file's raw data."""
    from scipy.io import netcdf_file
    from oceansar_amd import configs
    from oceansar_amd.sar_processor import FocusParams, sar_focus
    from oceansar_amd.sar_raw import sar_raw
    from oceansar_amd.cfg import ConfigFile
    text = configs.workload('small', L=64., n_pulses=1024, K=8, num_ch=2, b_ati=[0., 10.])
    cfg = configs.write(str(tmp_path / 'sim.cfg'), text)
    raw_p, slc_p = str(tmp_path / 'raw.nc'), str(tmp_path / 'slc.nc')
    sar_raw(cfg, raw_p, str(tmp_path / 'ocean.nc'), False, None, False, verbose=False)
    slc = sar_focus(cfg, raw_p, slc_p)
    with netcdf_file(raw_p, 'r', mmap=False) as f:
        raw = f.variables['raw_data_r'][:] + 1j * f.variables['raw_data_i'][:]
        sr0 = float(f.variables['sr0'].getValue())
    fp = FocusParams.from_cfg(ConfigFile(cfg), sr0)
    for ch in range(2):
        want = _oracle(raw[:, ch].astype(np.complex64), fp)
        e = rel_rms(slc[ch], want)
        print('sar_focus channel %d %s: rel RMS %.2e' % (ch, slc[ch].shape, e))
        assert e < 1e-12
    with netcdf_file(slc_p, 'r', mmap=False) as f:
        got = f.variables['slc_r'][:] + 1j * f.variables['slc_i'][:]
        assert f.variables['slc_r'].dimensions == ('ch_dim', 'pol_dim', 'az_dim', 'rg_dim')
    np.testing.assert_array_equal(got, slc)
