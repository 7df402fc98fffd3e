"""NetCDF files of the raw-data path and its consumers: RawFile,
OceanStateFile, ProcFile and L1bFile.

Same variable names, dimensions, units and the '_r'/'_i' convention for
complex values as oceansar/ocs_io/raw.py:160-250, state.py:7-72,
processed.py:6-77, insar_l1b.py:7-90 and netcdf.py:4-45, so downstream
readers (sar_processor.py:71-77, insar_processor.py:80-94,
ati_processor.py:78-96) find what they expect. netCDF4 is not a dependency: files are written by
scipy.io.netcdf_file as NetCDF-3 64-bit-offset, which netCDF4.Dataset reads.
NetCDF-3 has no 'u4', so the state file stores Nx/Ny as 'i4'.
"""
import time

import numpy as np
from scipy.io import netcdf_file


class _Handler:
    """get/set with a trailing '*' for complex variables (netcdf.py:20-45)."""

    def close(self):
        self._f.close()

    def get(self, name):
        v = self._f.variables
        if name[-1] == '*':
            re = v[name[:-1] + '_r'][:]
            c = np.empty(re.shape, dtype=np.complex64)
            c.real = re
            c.imag = v[name[:-1] + '_i'][:]
            return c
        var = v[name]
        return var.getValue() if var.shape == () else var[:].copy()

    def set(self, name, value):
        v = self._f.variables
        if name[-1] == '*':
            v[name[:-1] + '_r'][:] = np.real(value)
            v[name[:-1] + '_i'][:] = np.imag(value)
        elif v[name].shape == ():
            # scalar variable (assignValue indexes a 0-d array with [:] and fails on numpy 2)
            v[name].data[...] = value
        else:
            v[name][:] = value

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _var(f, name, typ, dims, units):
    x = f.createVariable(name, typ, dims)
    x.units = units
    return x


class RawFile(_Handler):
    """Raw data file (raw.py:160-250): raw_data_r/_i f8 [pol, ch, az, rg]."""

    def __init__(self, file_name, mode, raw_dim=None):
        self._f = netcdf_file(file_name, mode, version=2, mmap=False)
        if mode == 'w':
            f = self._f
            f.description = 'OCEANSAR SAR Raw Data File'
            f.history = 'Created ' + time.ctime(time.time())
            f.source = 'OCEANSAR Simulator (oceansar_amd)'
            if not raw_dim:
                raise ValueError('Raw data dimensions are needed when creating a new file!')
            if len(raw_dim) != 4:
                raise ValueError('raw_dim must be (pol, ch, az, rg)')
            for n, s in zip(('pol_dim', 'ch_dim', 'az_dim', 'rg_dim'), raw_dim):
                f.createDimension(n, int(s))
            _var(f, 'num_ch', 'i4', (), '[]')
            dims = ('pol_dim', 'ch_dim', 'az_dim', 'rg_dim')
            _var(f, 'raw_data_r', 'f8', dims, '[]')
            _var(f, 'raw_data_i', 'f8', dims, '[]')
            for n, u in (('inc_angle', '[deg]'), ('f0', '[Hz]'), ('ant_L', '[m]'),
                         ('prf', '[Hz]'), ('v_ground', '[m/s]'), ('orbit_alt', '[m]'),
                         ('sr0', '[m]'), ('rg_sampling', '[Hz]'), ('rg_bw', '[Hz]')):
                _var(f, n, 'f8', (), u)
            _var(f, 'b_ati', 'f8', ('ch_dim',), '[m]')
            _var(f, 'b_xti', 'f8', ('ch_dim',), '[m]')
            _var(f, 'NRCS_avg', 'f8', ('pol_dim', 'az_dim'), '[]')


class OceanStateFile(_Handler):
    """Surface reuse file (state.py:7-72): grid, wind, swell and wave_coefs_r/_i."""

    SCALARS = ('Lx', 'Ly', 'dx', 'dy', 'wind_dir', 'wind_dir_eff', 'wind_fetch', 'wind_U',
               'wind_U_eff', 'current_mag', 'current_dir', 'swell_enable', 'swell_ampl',
               'swell_dir', 'swell_wl', 'swell_ph0', 'choppy_enable')

    def __init__(self, file_name, mode, ocean_dim=None):
        self._f = netcdf_file(file_name, mode, version=2, mmap=False)
        if mode == 'w':
            f = self._f
            f.description = 'OCEANSAR Library Ocean State File'
            f.history = 'Created ' + time.ctime(time.time())
            f.source = 'OCEANSAR Library (oceansar_amd)'
            if not ocean_dim:
                raise ValueError('Ocean dimensions are needed when creating new file!')
            f.createDimension('y_dim', int(ocean_dim[0]))
            f.createDimension('x_dim', int(ocean_dim[1]))
            _var(f, 'Nx', 'i4', (), '[]')
            _var(f, 'Ny', 'i4', (), '[]')
            for n in self.SCALARS:
                _var(f, n, 'f4', (), '[]')
            _var(f, 'kx', 'f4', ('y_dim', 'x_dim'), '[rad/m]')
            _var(f, 'ky', 'f4', ('y_dim', 'x_dim'), '[rad/m]')
            _var(f, 'wave_coefs_r', 'f8', ('y_dim', 'x_dim'), '[]')
            _var(f, 'wave_coefs_i', 'f8', ('y_dim', 'x_dim'), '[]')


class ProcFile(_Handler):
    """Processed SLC file (oceansar/ocs_io/processed.py:6-77): slc_r/_i f8
    [ch, pol, az, rg], the radar scalars and the baselines."""

    def __init__(self, file_name, mode, proc_dim=None):
        self._f = netcdf_file(file_name, mode, version=2, mmap=False)
        if mode == 'w':
            f = self._f
            f.description = 'OCEANSAR Processed SLC Data File'
            f.history = 'Created ' + time.ctime(time.time())
            f.source = 'OCEANSAR Simulator (oceansar_amd)'
            if not proc_dim:
                raise ValueError('Processed raw data dimensions are needed when creating a new file!')
            if len(proc_dim) != 4:
                raise ValueError('proc_dim must be (ch, pol, az, rg)')
            for n, s in zip(('ch_dim', 'pol_dim', 'az_dim', 'rg_dim'), proc_dim):
                f.createDimension(n, int(s))
            _var(f, 'num_ch', 'i4', (), '[]')
            dims = ('ch_dim', 'pol_dim', 'az_dim', 'rg_dim')
            _var(f, 'slc_r', 'f8', dims, '[]')
            _var(f, 'slc_i', 'f8', dims, '[]')
            for n, u in (('inc_angle', '[deg]'), ('f0', '[Hz]'), ('ant_L', '[m]'),
                         ('prf', '[Hz]'), ('v_ground', '[m/s]'), ('orbit_alt', '[m]'),
                         ('sr0', '[m]'), ('rg_sampling', '[Hz]'), ('rg_bw', '[Hz]')):
                _var(f, n, 'f8', (), u)
            _var(f, 'b_ati', 'f8', ('ch_dim',), '[m]')
            _var(f, 'b_xti', 'f8', ('ch_dim',), '[m]')


class L1bFile(_Handler):
    """InSAR L1b file (oceansar/ocs_io/insar_l1b.py:7-90): ml_intensity f8
    [ch, pol, az, rg], ml_coherence / ml_phase f8 [covel, az, rg] with covel =
    C (C - 1) / 2 pairs of the C = ch * pol images, coh_lut i4 [ch, pol, ch,
    pol], the radar scalars, the baselines and the multilook window sizes."""

    def __init__(self, file_name, mode, data_dim=None):
        self._f = netcdf_file(file_name, mode, version=2, mmap=False)
        if mode == 'w':
            f = self._f
            f.description = 'OCEANSAR InSAR L1b data file'
            f.history = 'Created ' + time.ctime(time.time())
            f.source = 'OCEANSAR Simulator (oceansar_amd)'
            if not data_dim:
                raise ValueError('Processed raw data dimensions are needed when creating a new file!')
            if len(data_dim) != 4:
                raise ValueError('data_dim must be (ch, pol, az, rg)')
            for n, s in zip(('ch_dim', 'pol_dim', 'az_dim', 'rg_dim'), data_dim):
                f.createDimension(n, int(s))
            tot = int(data_dim[0]) * int(data_dim[1])
            f.createDimension('covel_dim', tot * (tot - 1) // 2)
            _var(f, 'num_ch', 'i4', (), '[]')
            _var(f, 'ml_intensity', 'f8', ('ch_dim', 'pol_dim', 'az_dim', 'rg_dim'), 'm2/m2')
            _var(f, 'ml_coherence', 'f8', ('covel_dim', 'az_dim', 'rg_dim'), '[]')
            _var(f, 'ml_phase', 'f8', ('covel_dim', 'az_dim', 'rg_dim'), 'rad')
            lut = _var(f, 'coh_lut', 'i4', ('ch_dim', 'pol_dim', 'ch_dim', 'pol_dim'), '[]')
            lut.description = ('LUT that indicates where to find coh and phase for any '
                               'combination in ml_coh and ml_phase')
            for n, u in (('inc_angle', '[deg]'), ('f0', '[Hz]'), ('az_sampling', '[Hz]'),
                         ('v_ground', '[m/s]'), ('orbit_alt', '[m]'), ('sr0', '[m]'),
                         ('rg_sampling', '[Hz]')):
                _var(f, n, 'f8', (), u)
            _var(f, 'b_ati', 'f8', ('ch_dim',), '[m]')
            _var(f, 'b_xti', 'f8', ('ch_dim',), '[m]')
            _var(f, 'rg_ml', 'i4', (), '[]')
            _var(f, 'az_ml', 'i4', (), '[]')
