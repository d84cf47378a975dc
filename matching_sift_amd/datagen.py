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


def pairs_batch(n_pairs, seed=30000, sizes=None, frac_copy=0.4, noise_sd=8.0):
    """Batched image pairs: pair p has a base image of sizes[p,1] descriptors
    (as base()) and a query image of sizes[p,0] descriptors, frac_copy of them
    noisy copies of the base image's rows.  sizes default to pair_sizes().
    Returns (q[Qtot,128], q_off[P+1], r[Rtot,128], r_off[P+1], pairs[P,2])
    with pair p = (query image p, base image p)."""
    if sizes is None:
        sizes = pair_sizes(n_pairs, seed=seed)
    sizes = np.asarray(sizes, np.int64).reshape(n_pairs, 2)
    q_off = np.zeros(n_pairs + 1, np.uint32)
    r_off = np.zeros(n_pairs + 1, np.uint32)
    q_off[1:] = np.cumsum(sizes[:, 0])
    r_off[1:] = np.cumsum(sizes[:, 1])
    q = np.empty((int(q_off[-1]), 128), np.float32)
    r = np.empty((int(r_off[-1]), 128), np.float32)
    for p in range(n_pairs):
        nq, nr = int(sizes[p, 0]), int(sizes[p, 1])
        rb = base(nr, seed=seed + 1 + 2 * p)
        r[r_off[p]:r_off[p + 1]] = rb
        if nr > 0:
            q[q_off[p]:q_off[p + 1]] = queries(rb, nq, seed=seed + 2 + 2 * p, frac_copy=frac_copy,
                                               noise_sd=noise_sd)
        else:
            q[q_off[p]:q_off[p + 1]] = base(nq, seed=seed + 2 + 2 * p)
    pairs = np.stack([np.arange(n_pairs), np.arange(n_pairs)], 1).astype(np.uint32)
    return q, q_off, r, r_off, pairs


def config3(n_pairs=512):
    """512 image pairs, keypoints per image LogNormal(median 2000, sigma 0.6)
    clipped to [200, 8000], 40% of each query image noisy copies."""
    return pairs_batch(n_pairs)


def sha256(arr):
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()
