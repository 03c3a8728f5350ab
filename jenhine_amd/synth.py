"""Synthetic scan data: seeded ellipsoid phantoms, analytic line integrals,
gain/dark frames and raw uint16 counts (binding of libjhsynth.so)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
EREC = 16


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libjhsynth.so")
        if not os.path.exists(path):
            from . import build
            build.build_synth()
        L = ctypes.CDLL(path)
        L.jh_synth_phantom.restype = ctypes.c_int
        L.jh_synth_phantom.argtypes = [ctypes.c_uint64, ctypes.c_int] + [ctypes.c_double] * 8 + \
            [ctypes.c_int, ctypes.c_double, ctypes.c_void_p]
        L.jh_synth_lineint.restype = ctypes.c_int
        L.jh_synth_lineint.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.jh_synth_voxelize.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L.jh_synth_gain_dark.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.jh_synth_raw.argtypes = [ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.jh_synth_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_void_p]
        L.jh_synth_normal.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_void_p]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def phantom(seed, n, half_extent, frac=0.6, axes=(2.0, 12.0), mu=(0.01, 0.05), tooth=False, tooth_mu=0.3):
    """Ellipsoid records (n[+1] x 16: centre, semi-axes, rotation, mu)."""
    hx, hy, hz = half_extent
    out = np.zeros((n + (1 if tooth else 0), EREC), np.float64)
    m = _lib().jh_synth_phantom(seed, n, hx, hy, hz, frac, axes[0], axes[1], mu[0], mu[1], int(tooth),
                                tooth_mu, _p(out))
    assert m == out.shape[0]
    return out


def line_integrals(ptm, n_u, n_v, ellipsoids):
    ptm = np.ascontiguousarray(ptm, np.float64).reshape(-1, 12)
    ell = np.ascontiguousarray(ellipsoids, np.float64)
    out = np.zeros((ptm.shape[0], n_v, n_u), np.float32)
    rc = _lib().jh_synth_lineint(_p(ptm), ptm.shape[0], n_u, n_v, _p(ell), ell.shape[0], _p(out))
    if rc != 0:
        raise ValueError("singular projection matrix")
    return out


def voxelize(ellipsoids, shape_zyx, pxy, pz):
    Z, Y, X = shape_zyx
    ell = np.ascontiguousarray(ellipsoids, np.float64)
    out = np.zeros((Z, Y, X), np.float32)
    _lib().jh_synth_voxelize(_p(ell), ell.shape[0], X, Y, Z, pxy, pz, _p(out))
    return out


def gain_dark(seed, n_u, n_v, g=(0.9, 1.1), d=(90, 110)):
    gain = np.zeros((n_v, n_u), np.float32)
    dark = np.zeros((n_v, n_u), np.uint16)
    _lib().jh_synth_gain_dark(seed, n_u * n_v, g[0], g[1], d[0], d[1], _p(gain), _p(dark))
    return gain, dark


def raw_counts(seed, lineint, I0, gain=None, dark=None, noise=True):
    p = np.ascontiguousarray(lineint, np.float32)
    N, V, U = p.shape
    out = np.zeros((N, V, U), np.uint16)
    g = None if gain is None else np.ascontiguousarray(gain, np.float32)
    d = None if dark is None else np.ascontiguousarray(dark, np.uint16)
    _lib().jh_synth_raw(seed, _p(p), N, U * V, I0, _p(g), _p(d), int(noise), _p(out))
    return out


def uniform(seed, shape, lo=0.0, hi=1.0, stream=0):
    out = np.zeros(shape, np.float32)
    _lib().jh_synth_uniform(seed, stream, out.size, lo, hi, _p(out))
    return out


def normal(seed, shape, stream=0):
    out = np.zeros(shape, np.float32)
    _lib().jh_synth_normal(seed, stream, out.size, _p(out))
    return out
