"""oceansar_amd: MI355X-native OCEANSAR raw-data generation hot path.

Host side in Python (cfg, geometry, surface initialisation, the sar_raw
driver), per-pulse work in hand-written HIP kernels for gfx950 reached
through a ctypes C-ABI (liboceansar_hip.so).
"""
__version__ = "0.1.0"
