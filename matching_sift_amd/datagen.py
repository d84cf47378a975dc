"""Seeded synthetic workloads (binding of libsiftdata.so, csrc/datagen.c).

The distributions follow the benchmark configurations: base descriptors
round(Exponential(mean 20)) with 30% zeros, clipped to 255; queries are 50%
noisy copies (Normal(0, 8) noise, rounded, clipped) of random base rows and 50%
fresh draws.
"""
import ctypes as C
import hashlib
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_PKG, "libsiftdata.so")
_L = None

SEED_BASE = 20221
SEED_QUERY = 20222


def _lib():
    global _L
    if _L is None:
        if not os.path.exists(_PATH):
            raise OSError("libsiftdata.so not built; run __graft_entry__.build()")
        L = C.CDLL(_PATH)
        L.sift_gen_base.argtypes = [C.c_uint64, C.c_int64, C.c_void_p]
        L.sift_gen_queries.argtypes = [C.c_uint64, C.c_void_p, C.c_int64, C.c_int64, C.c_double,
                                       C.c_double, C.c_void_p, C.c_void_p]
        L.sift_gen_pair_sizes.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double,
                                          C.c_int64, C.c_int64, C.c_void_p]
        L.sift_gen_unit.argtypes = [C.c_uint64, C.c_int64, C.c_void_p]
        L.sift_gen_gauss_int.argtypes = [C.c_uint64, C.c_int64, C.c_double, C.c_double, C.c_void_p]
        L.sift_gen_uniform.argtypes = [C.c_uint64, C.c_int64, C.c_void_p]
        for f in (L.sift_gen_base, L.sift_gen_queries, L.sift_gen_pair_sizes, L.sift_gen_unit,
                  L.sift_gen_gauss_int, L.sift_gen_uniform):
            f.restype = None
        _L = L
    return _L


def base(n, seed=SEED_BASE):
    out = np.empty((n, 128), np.float32)
    _lib().sift_gen_base(seed, n, out.ctypes.data)
    return out


def queries(base_arr, m, seed=SEED_QUERY, frac_copy=0.5, noise_sd=8.0, return_src=False):
    out = np.empty((m, 128), np.float32)
    src = np.empty((m,), np.int64)
    b = np.ascontiguousarray(base_arr, np.float32)
    _lib().sift_gen_queries(seed, b.ctypes.data, b.shape[0], m, frac_copy, noise_sd,
                            out.ctypes.data, src.ctypes.data)
    return (out, src) if return_src else out


def pair_sizes(pairs, seed=7, median=2000.0, sigma=0.6, lo=200, hi=8000):
    out = np.empty((pairs, 2), np.int64)
    _lib().sift_gen_pair_sizes(seed, pairs, median, sigma, lo, hi, out.ctypes.data)
    return out


def unit(n, seed=11):
    out = np.empty((n, 128), np.float32)
    _lib().sift_gen_unit(seed, n, out.ctypes.data)
    return out


def gauss_int(n, sd, lim, seed=13):
    out = np.empty((n,), np.float32)
    _lib().sift_gen_gauss_int(seed, n, sd, lim, out.ctypes.data)
    return out


def uniform(n, seed=17):
    out = np.empty((n,), np.float64)
    _lib().sift_gen_uniform(seed, n, out.ctypes.data)
    return out


def config0():
    """N=10,000 base, M=1,000 queries (the CPU-runnable exact configuration)."""
    b = base(10_000)
    return queries(b, 1_000), b


def config1(n=1_000_000, m=10_000):
    """N=1,000,000 base, M=10,000 queries."""
    b = base(n)
    return queries(b, m), b


def sha256(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()
