"""ctypes binding of the oracle's C restatement (test infrastructure only)."""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            from . import build as _build
            _build.build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        u64 = ctypes.c_uint64
        L.oracle_philox4x32_10.argtypes = [P, u64, u64, P]
        L.oracle_keyed_uniforms.argtypes = [P, i64, P, i64, P, i64, u64, u64,
                                            u64, P, P]
        L.oracle_levy_f32.argtypes = [ctypes.c_double, P, P, P, i64]
        L.oracle_levy_f64.argtypes = [ctypes.c_double, P, P, P, i64]
        L.oracle_keyed_levy.argtypes = [P, i64, P, i64, P, i64, u64, u64, u64,
                                        ctypes.c_double, ctypes.c_int, P, P]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Ten Philox4x32 rounds; restates steam/noise.py:67-85."""
    inp = np.array([int(c0) & (2**64 - 1), int(c1) & (2**64 - 1),
                    int(c2) & (2**64 - 1), int(c3) & (2**64 - 1)],
                   dtype=np.uint64)
    out = np.zeros(4, dtype=np.uint64)
    lib().oracle_philox4x32_10(_ptr(inp), int(k0), int(k1), _ptr(out))
    return tuple(int(v) for v in out)


def keyed_uniforms(world_ix, world_iy, world_iz, key0, key1, key2, lane0,
                   lane1):
    """Fill two float32 lanes in C order; restates steam/noise.py:103-130."""
    ix = np.ascontiguousarray(world_ix, dtype=np.int64)
    iy = np.ascontiguousarray(world_iy, dtype=np.int64)
    iz = np.ascontiguousarray(world_iz, dtype=np.int64)
    n = ix.size * iy.size * iz.size
    assert lane0.dtype == np.float32 and lane0.size == n and lane0.flags.c_contiguous
    assert lane1.dtype == np.float32 and lane1.size == n and lane1.flags.c_contiguous
    lib().oracle_keyed_uniforms(_ptr(ix), ix.size, _ptr(iy), iy.size, _ptr(iz),
                                iz.size, int(key0), int(key1), int(key2),
                                _ptr(lane0), _ptr(lane1))


def levy_f32(alpha, u0, u1):
    """float32 CMS with the reference op order (steam/simulate.py:413-479)."""
    u0 = np.ascontiguousarray(u0, dtype=np.float32)
    u1 = np.ascontiguousarray(u1, dtype=np.float32)
    out = np.empty(u0.shape, dtype=np.float32)
    lib().oracle_levy_f32(float(alpha), _ptr(u0), _ptr(u1), _ptr(out), u0.size)
    return out


def levy_f64(alpha, u0, u1):
    """float64 CMS on the same uniforms (the FP64 oracle)."""
    u0 = np.ascontiguousarray(u0, dtype=np.float32)
    u1 = np.ascontiguousarray(u1, dtype=np.float32)
    out = np.empty(u0.shape, dtype=np.float64)
    lib().oracle_levy_f64(float(alpha), _ptr(u0), _ptr(u1), _ptr(out), u0.size)
    return out


def keyed_levy(world_ix, world_iy, world_iz, key, alpha, dtype=np.float64):
    """Keyed Levy draws over an index box of centres, C order."""
    ix = np.ascontiguousarray(world_ix, dtype=np.int64)
    iy = np.ascontiguousarray(world_iy, dtype=np.int64)
    iz = np.ascontiguousarray(world_iz, dtype=np.int64)
    shape = (ix.size, iy.size, iz.size)
    if dtype == np.float32:
        out = np.empty(shape, dtype=np.float32)
        lib().oracle_keyed_levy(_ptr(ix), ix.size, _ptr(iy), iy.size, _ptr(iz),
                                iz.size, int(key[0]), int(key[1]), int(key[2]),
                                float(alpha), 0, _ptr(out), None)
    else:
        out = np.empty(shape, dtype=np.float64)
        lib().oracle_keyed_levy(_ptr(ix), ix.size, _ptr(iy), iy.size, _ptr(iz),
                                iz.size, int(key[0]), int(key[1]), int(key[2]),
                                float(alpha), 1, None, _ptr(out))
    return out
