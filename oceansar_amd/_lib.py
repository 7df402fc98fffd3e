"""ctypes binding of liboceansar_hip.so (include/oceansar_hip.h).

The library is built in-tree by ``oceansar_amd._build`` (or
``__graft_entry__.build()``). Loading fails loudly when it is missing: there
is no CPU fallback on the product path. Device buffers are torch tensors on
a ROCm device; only their data pointers cross the boundary.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_void_p

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'liboceansar_hip.so')
MAX_CH = 8

ECHO_BRAGG = 1
ECHO_XTI = 2
ECHO_SPECULAR = 4


class OsarError(RuntimeError):
    """Raised when a library call returns a negative status."""

    def __init__(self, code, msg):
        super().__init__('%s (code %d)' % (msg, code))
        self.code = code


class EchoDesc(ctypes.Structure):
    _fields_ = [(n, c_int32) for n in (
        'nx', 'ny', 'n_pulses_total', 'pulse0', 'npulses', 'nbins', 'n_ch', 'n_pol', 'K',
        'snap0', 'n_snap_slots', 'rcs0', 'n_rcs_slots', 'flags', 'pulses_per_block',
        'rows_per_block', 'window_bins', 'pol', 'cols_per_block', 'scat_len')] + \
        [(n, c_double) for n in (
            't_step', 't_span', 'v_ground', 'sr0', 'gr0', 'k0', 'min_sr', 'binsize',
            'ltx_over_l', 'lrx_over_l', 'sr_col_min', 'sr_col_max', 'resid_max')] + \
        [('b_ati', c_double * MAX_CH), ('b_xti', c_double * MAX_CH)] + \
        [(n, c_void_p) for n in (
            'sr_col', 'sin_inc', 'cos_inc', 'dinc_col', 'cap_rate', 'y_row', 'dx', 'dy', 'dz',
            'amp', 'rcs_slot', 'scat_m', 'scat_p', 'ind0_m', 'ind0_p', 'sinc_tab', 'raw',
            'sx', 'sy', 'sxx', 'syy', 'sxy', 'inc_col')] + \
        [('facet_dx', c_double), ('facet_dy', c_double), ('spec_energy', c_void_p)] + \
        [(n, c_void_p) for n in ('amp64', 'scat64_m', 'scat64_p', 'raw64')]


class BraggDesc(ctypes.Structure):
    _fields_ = [(n, c_int32) for n in ('nx', 'ny', 'pol', 'spread', 'n_spec', 'pulse')] + \
        [(n, c_double) for n in ('k0', 'd', 'wind_dir', 'wind_U', 'wind_fetch', 'dx', 'dy',
                                 't_step', 't_span', 'v_ground', 'gr0')] + \
        [(n, c_int32) for n in ('spec_uniform', 'fp64')] + \
        [(n, c_double) for n in ('spec_k0', 'spec_dk', 'spec_inv_dk')] + \
        [(n, c_void_p) for n in ('cos_inc', 'sin_inc', 'y_row', 'spec_k', 'spec_S', 'spec_slope',
                                 'diffx', 'diffy', 'amp', 'rcs_sum')] + \
        [(n, c_int32) for n in ('n_refresh', 'snap0')] + \
        [(n, c_void_p) for n in ('refresh_pulse', 'refresh_snap')] + \
        [(n, ctypes.c_int64) for n in ('diff_stride', 'amp_stride')] + \
        [('amp64', c_void_p), ('hmtf', c_void_p)]


NUM_FIELDS = 15   # OSAR_NUM_FIELDS: output slots of osar_surface_desc


class SurfaceDesc(ctypes.Structure):
    _fields_ = [(n, c_int32) for n in ('nx', 'ny', 'fields', 'choppy', 'n_times',
                                       'mirror_grid', 'slope_fp32', 'reserved')] + \
        [(n, c_double) for n in ('depth', 'current_x', 'current_y', 'kx_res', 'ky_res')] + \
        [(n, c_void_p) for n in ('kx', 'ky', 'coefs', 'times')] + \
        [('out', c_void_p * NUM_FIELDS)]


class EchoBatchDesc(ctypes.Structure):
    _fields_ = [('n', c_int32), ('workgroups', c_int32), ('descs', c_void_p),
                ('plan_host', c_void_p), ('plan_dev', c_void_p)]


ECHO_BATCH_MAX = 4096   # OSAR_ECHO_BATCH_MAX


class InitDesc(ctypes.Structure):
    _fields_ = [(n, c_int32) for n in ('nx', 'ny', 'spread', 'swell')] + \
        [('seed', ctypes.c_uint64)] + \
        [(n, ctypes.c_float) for n in ('kx_res', 'ky_res', 'wind_dir_eff', 'reserved')] + \
        [(n, c_double) for n in ('k_p', 'c_p', 'omega_w', 'alpha_p', 'sigma', 'gamma',
                                 'alpha_m', 'c_m', 'k_m', 'a_0', 'a_p', 'a_m', 'cK_p', 'cK_m',
                                 'swell_dir', 'freq_r', 'sigf', 'sigs', 'swell_amp')] + \
        [(n, c_void_p) for n in ('kx', 'ky', 'coefs', 'dirspec', 'rmss_sums')]


class FocusDesc(ctypes.Structure):
    _fields_ = [(n, c_int32) for n in ('n_pol', 'n_az0', 'n_rg0', 'n_az', 'n_rg', 'az_lo',
                                       'az_hi', 'reserved')] + \
        [('raw_pol_stride', c_int64)] + \
        [(n, c_void_p) for n in ('raw', 'fr', 'rcmc', 'h', 'work', 'slc')]


class CoregDesc(ctypes.Structure):
    _fields_ = [(n, c_int32) for n in ('n_pol', 'n_az', 'n_rg', 'n_fft')] + \
        [('in_pol_stride', c_int64)] + \
        [(n, c_void_p) for n in ('in_', 'kernel', 'flat', 'out')]


class MultilookDesc(ctypes.Structure):
    _fields_ = [(n, c_int32) for n in ('n_img', 'n_az', 'n_rg', 'is_complex', 'rg_len',
                                       'az_len', 'rg_flat', 'az_flat', 'az_lo', 'az_hi',
                                       'rg_lo', 'rg_hi')] + \
        [('img_stride', c_int64)] + \
        [(n, c_void_p) for n in ('img', 'rg_taps', 'az_taps', 'work', 'intensity', 'coherence',
                                 'phase')]


FIELD_ORDER = ['Dx', 'Dy', 'Dz', 'Diffx', 'Diffy', 'Diffxx', 'Diffyy', 'Diffxy',
               'Vx', 'Vy', 'Vz', 'Ax', 'Ay', 'Az', 'hMTF']
FIELD_BITS = {n: 1 << i for i, n in enumerate(FIELD_ORDER)}

EXPORTS = ['osar_ctx_create', 'osar_ctx_destroy', 'osar_last_error', 'osar_version',
           'osar_struct_size', 'osar_echo_accumulate', 'osar_echo_accumulate_fp64',
           'osar_fill_zero_c32',
           'osar_echo_deposit', 'osar_facet_bragg', 'osar_fill_zero_f64',
           'osar_surface_workspace_bytes', 'osar_surface_synth', 'osar_surface_synth_noalloc',
           'osar_sar_focus', 'osar_surface_init', 'osar_echo_batch_plan_bytes',
           'osar_echo_accumulate_batch', 'osar_insar_coregister', 'osar_multilook_work_bytes',
           'osar_insar_multilook']
SYNTH_SCRATCH_BYTES = 256 << 20    # OSAR_SYNTH_SCRATCH_BYTES

_lib = None


def load():
    """Load the library once and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError('liboceansar_hip.so is not built (%s); run '
                          'python -m oceansar_amd._build' % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    lib.osar_ctx_create.argtypes = [c_int, POINTER(c_void_p)]
    lib.osar_ctx_destroy.argtypes = [c_void_p]
    lib.osar_last_error.restype = ctypes.c_char_p
    lib.osar_version.restype = c_int
    lib.osar_struct_size.argtypes = [c_int]
    lib.osar_struct_size.restype = c_int64
    lib.osar_echo_accumulate.argtypes = [c_void_p, POINTER(EchoDesc), c_void_p]
    lib.osar_echo_accumulate_fp64.argtypes = [c_void_p, POINTER(EchoDesc), c_void_p]
    lib.osar_fill_zero_c32.argtypes = [c_void_p, c_void_p, c_int64, c_void_p]
    lib.osar_fill_zero_f64.argtypes = [c_void_p, c_void_p, c_int64, c_void_p]
    lib.osar_echo_deposit.argtypes = [c_void_p, c_void_p, c_void_p, c_int64, c_double, c_double,
                                      c_void_p, c_int32, c_int32, c_void_p, c_int32, c_void_p]
    lib.osar_facet_bragg.argtypes = [c_void_p, POINTER(BraggDesc), c_void_p]
    lib.osar_surface_workspace_bytes.argtypes = [POINTER(SurfaceDesc)]
    lib.osar_surface_workspace_bytes.restype = c_int64
    lib.osar_surface_synth.argtypes = [c_void_p, POINTER(SurfaceDesc), c_void_p]
    lib.osar_surface_synth_noalloc.argtypes = [c_void_p, POINTER(SurfaceDesc), c_void_p]
    lib.osar_sar_focus.argtypes = [c_void_p, POINTER(FocusDesc), c_void_p]
    lib.osar_surface_init.argtypes = [c_void_p, POINTER(InitDesc), c_void_p]
    lib.osar_echo_batch_plan_bytes.argtypes = [c_int32]
    lib.osar_echo_batch_plan_bytes.restype = c_int64
    lib.osar_echo_accumulate_batch.argtypes = [c_void_p, POINTER(EchoBatchDesc), c_void_p]
    lib.osar_insar_coregister.argtypes = [c_void_p, POINTER(CoregDesc), c_void_p]
    lib.osar_multilook_work_bytes.argtypes = [POINTER(MultilookDesc)]
    lib.osar_multilook_work_bytes.restype = c_int64
    lib.osar_insar_multilook.argtypes = [c_void_p, POINTER(MultilookDesc), c_void_p]
    for name, cls in ((0, EchoDesc), (1, BraggDesc), (2, SurfaceDesc), (3, FocusDesc),
                      (4, InitDesc), (5, CoregDesc), (6, MultilookDesc)):
        n = lib.osar_struct_size(name)
        if n != ctypes.sizeof(cls):
            raise ImportError('descriptor %s: C size %d != ctypes size %d'
                              % (cls.__name__, n, ctypes.sizeof(cls)))
    _lib = lib
    return lib


def check(code):
    if code != 0:
        msg = load().osar_last_error().decode(errors='replace')
        raise OsarError(code, msg)


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    return None if t is None else c_void_p(t.data_ptr())


def stream_ptr(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return c_void_p(s.cuda_stream)


class Context:
    """Owns one osar_ctx for a device."""

    def __init__(self, device=0):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('oceansar_amd needs a ROCm GPU; none is visible')
        self.lib = load()
        self.device = device
        h = c_void_p()
        check(self.lib.osar_ctx_create(device, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.osar_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ctx = {}


def context(device=0):
    if device not in _ctx:
        _ctx[device] = Context(device)
    return _ctx[device]


def echo_deposit(srg, scene, binsize, min_sr, sinc_vec, n_sinc_samples, sinc_ovs, output,
                 rg_only=False):
    """chan_profile_numba on the device (range_profile.py:83-116)."""
    import torch
    if rg_only:
        raise NotImplementedError('rg_only (profile_integrator_1d) is not on the raw path')
    if srg.dtype != torch.float64 or not srg.is_cuda:
        raise TypeError('srg must be a float64 device tensor')
    if scene.dtype == torch.complex128:
        scene = scene.to(torch.complex64)
    if scene.dtype != torch.complex64 or not scene.is_cuda:
        raise TypeError('scene must be a complex device tensor')
    if output.dtype != torch.complex64 or not output.is_cuda or not output.is_contiguous():
        raise TypeError('output must be a contiguous complex64 device tensor')
    if srg.numel() != scene.numel():
        raise TypeError('srg and scene sizes differ')
    if not torch.is_tensor(sinc_vec):
        sinc_vec = torch.as_tensor(sinc_vec, dtype=torch.float32, device=srg.device)
    srg = srg.contiguous()
    scene = scene.contiguous()
    ctx = context(srg.device.index or 0)
    check(ctx.lib.osar_echo_deposit(ctx.h, ptr(srg), ptr(scene), srg.numel(), float(binsize),
                                    float(min_sr), ptr(sinc_vec), int(n_sinc_samples),
                                    int(sinc_ovs), ptr(output), output.numel(), stream_ptr()))
