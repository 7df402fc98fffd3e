"""Golden vectors of the reference's sr_to_geo, the one routine of the InSAR
flat-earth step that no closed form pins to the last bits (its quadratic
spline over a 2500-point look-angle table).

Run in a container where the reference checkout exists (read-only):
    python tests/golden/make_golden_insar.py /root/reference

oceansar/utils/geometry.py is executed from its source text in a fresh
namespace (never through the import system, so no bytecode is read or
written), with the reference's constants and utils.misc supplied the same way
(tests/golden/make_golden.py does the same for its leaves). Inputs: the slant
ranges of the range lines of the three workloads (sr0 + n c / (2 fs)); output
golden_insar.npz, data only (loaded with allow_pickle=False).
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from make_golden import _exec_source  # noqa: E402


def main(ref):
    import types
    sys.dont_write_bytecode = True
    pkg = os.path.join(ref, 'oceansar')
    const = _exec_source(os.path.join(pkg, 'constants', 'constants.py'), 'oceansar.constants')
    oceansar = types.ModuleType('oceansar')
    oceansar.constants = const
    sys.modules['oceansar'] = oceansar
    sys.modules['oceansar.constants'] = const
    misc = _exec_source(os.path.join(pkg, 'utils', 'misc.py'), 'oceansar.utils.misc')
    utils = types.ModuleType('oceansar.utils')
    utils.optimize_fftsize = misc.optimize_fftsize
    oceansar.utils = utils
    sys.modules['oceansar.utils'] = utils
    geom = _exec_source(os.path.join(pkg, 'utils', 'geometry.py'), 'oceansar.utils.geometry')
    g = {}
    alt = 514e3
    sr0 = geom.inc_to_sr(np.deg2rad(35.), alt)
    fs = 110e6
    for name, n in (('115', 115), ('267', 267), ('624', 624)):
        sr = sr0 + np.arange(n) / fs * const.c / 2
        gr, inc, look, b = geom.sr_to_geo(sr, alt)
        g['srgeo_sr_' + name] = sr
        g['srgeo_gr_' + name] = gr
        g['srgeo_inc_' + name] = inc
        g['srgeo_look_' + name] = look
        g['srgeo_b_' + name] = b
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden_insar.npz')
    np.savez(out, **g)
    print('wrote', out, len(g), 'arrays')


if __name__ == '__main__':
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
