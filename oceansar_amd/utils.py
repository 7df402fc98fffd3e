"""Sizing helpers (FFT-friendly sizes, dB conversion).

Restates oceansar/utils/misc.py:38-83 (factorize, nearest_power2,
optimize_fftsize) and misc.py:209-214 (db2lin) with integer arithmetic.
"""
import numpy as np


def factorize(n):
    """Prime factors of n in ascending order (oceansar/utils/misc.py:38-53)."""
    n = int(n)
    out = []
    p = 2
    while n > 1 and p * p <= n:
        while n % p == 0:
            out.append(p)
            n //= p
        p += 1 if p == 2 else 2
    if n > 1:
        out.append(n)
    return np.array(out, dtype=float)


def nearest_power2(values):
    """Smallest power of two >= value (oceansar/utils/misc.py:56-57)."""
    return (2 ** np.ceil(np.log2(values))).astype(int)


def _max_prime(n):
    f = factorize(n)
    return int(f.max()) if f.size else 1


def optimize_fftsize(values, max_prime=2):
    """Next size whose largest prime factor is <= max_prime.

    oceansar/utils/misc.py:60-83: max_prime == 2 rounds up to a power of two,
    otherwise the value is incremented until it is max_prime-smooth.
    Scalars return a scalar, arrays an array.
    """
    scalar = np.isscalar(values)
    vals = np.atleast_1d(np.asarray(values, dtype=int))
    if max_prime == 2:
        good = nearest_power2(vals)
    else:
        good = []
        for v in vals:
            b = int(v)
            while _max_prime(b) > max_prime:
                b += 1
            good.append(b)
        good = np.array(good, dtype=int)
    return int(good[0]) if (scalar or good.size == 1) else good


def db2lin(a, amplitude=False):
    """dB to linear (oceansar/utils/misc.py:209-214)."""
    return 10 ** (a / 20) if amplitude else 10 ** (a / 10)
