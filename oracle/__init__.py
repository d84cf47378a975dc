"""CPU oracle of the reference's 2-NN path (ctypes binding of liboracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this package, and only as the checker
(or the timed CPU baseline).  The product (matching_sift_amd) never uses it.

Function -> reference:
  host_brute        src/knn_brute.cu:801-835 (+ host_lenght :873-885)
  host_lenght       src/knn_brute.cu:873-885 (listed query/base pairs)
  host_sort         src/knn_brute.cu:839-870
  ratio             src/knn_brute.cu:430-442 (true L2, no unit-norm +2)
  lsh_codes         src/lsh_h.cu:265-317 (bit b set when h_b.x < 0)
  lsh_sort          src/lsh_h.cu:587-591 (stable by (code, index))
  lsh_query         src/lsh_h.cu:598-620, :657-681 (same-bucket candidates)
  lsh_quantize,     float mode of the index (data not all integers in [0,255]):
  lsh_scale_exp,    the integers hashed for x, and the re-rank by host_lenght
  lsh_query_f32
  mlp_forward       learned hash (no reference counterpart)
"""
import ctypes as C
import os
import subprocess

import numpy as np

_DIR = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(_DIR, "liboracle.so")
_L = None


def build():
    subprocess.run(["make", "-s", "-C", _DIR, "-B", "liboracle.so"], check=True)


def _lib():
    global _L
    if _L is None:
        if not os.path.exists(_PATH):
            build()
        L = C.CDLL(_PATH)
        vp, i32, f32 = C.c_void_p, C.c_int, C.c_float
        sig = {
            "oracle_host_brute": [vp, vp, i32, i32, vp],
            "oracle_host_lenght": [vp, vp, vp, vp, i32, vp],
            "oracle_host_brute_mt": [vp, vp, i32, i32, vp, i32],
            "oracle_host_brute_mutant": [vp, vp, i32, i32, vp, i32],
            "oracle_host_sort": [vp, i32, i32, vp],
            "oracle_knn2_int": [vp, vp, i32, i32, vp, vp, i32],
            "oracle_ratio": [vp, i32, f32, vp],
            "oracle_lsh_codes": [vp, i32, vp, i32, vp],
            "oracle_lsh_sort": [vp, i32, vp],
            "oracle_lsh_query": [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp, i32],
            "oracle_mlp_forward": [vp, i32, vp, vp, i32, vp, vp, i32, vp, vp],
            "oracle_hash_forward": [vp, i32, vp, i32, i32, vp, vp],
            "oracle_hash_backward": [vp, i32, vp, i32, i32, vp, vp, vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = None
        L.oracle_lsh_quantize.argtypes = [vp, C.c_size_t, i32, vp]
        L.oracle_lsh_quantize.restype = None
        L.oracle_lsh_scale_exp.argtypes = [vp, C.c_size_t]
        L.oracle_lsh_scale_exp.restype = i32
        L.oracle_lsh_query_f32.argtypes = [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp, i32]
        L.oracle_lsh_query_f32.restype = None
        L.oracle_hash_loss.argtypes = [vp, vp, i32, i32, C.c_double, C.c_double, C.c_double]
        L.oracle_hash_loss.restype = C.c_double
        L.oracle_hash_loss_grad.argtypes = [vp, i32, i32, C.c_double, C.c_double, C.c_double, vp, vp]
        L.oracle_hash_loss_grad.restype = C.c_double
        _L = L
    return _L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def host_brute(q, r, threads=1):
    """float32 [M,4] = {d1, d2, idx1, idx2} exactly as host_brute computes it."""
    q, r = _f32(q), _f32(r)
    out = np.empty((q.shape[0], 4), np.float32)
    if threads <= 1:
        _lib().oracle_host_brute(q.ctypes.data, r.ctypes.data, q.shape[0], r.shape[0], out.ctypes.data)
    else:
        _lib().oracle_host_brute_mt(q.ctypes.data, r.ctypes.data, q.shape[0], r.shape[0],
                                    out.ctypes.data, threads)
    return out


def host_lenght(q, r, qi, ri):
    """float32 [n]: host_lenght(q[qi[k]], r[ri[k]]) (src/knn_brute.cu:873-885)."""
    q, r = _f32(q), _f32(r)
    qi = np.ascontiguousarray(qi, np.int64)
    ri = np.ascontiguousarray(ri, np.int64)
    assert qi.shape == ri.shape and qi.ndim == 1
    assert (qi >= 0).all() and (qi < q.shape[0]).all() and (ri >= 0).all() and (ri < r.shape[0]).all()
    out = np.empty((qi.shape[0],), np.float32)
    _lib().oracle_host_lenght(q.ctypes.data, r.ctypes.data, qi.ctypes.data, ri.ctypes.data, qi.shape[0],
                              out.ctypes.data)
    return out


def host_brute_mutant(q, r, mutation):
    q, r = _f32(q), _f32(r)
    out = np.empty((q.shape[0], 4), np.float32)
    _lib().oracle_host_brute_mutant(q.ctypes.data, r.ctypes.data, q.shape[0], r.shape[0],
                                    out.ctypes.data, mutation)
    return out


def host_sort(dist):
    dist = _f32(dist)
    out = np.empty((dist.shape[0], 4), np.float32)
    _lib().oracle_host_sort(dist.ctypes.data, dist.shape[1], dist.shape[0], out.ctypes.data)
    return out


def knn2_int(q, r, threads=1):
    """(idx uint32[M,2], dist uint32[M,2]) in int64 arithmetic; UINT32_MAX = absent."""
    q, r = _f32(q), _f32(r)
    idx = np.empty((q.shape[0], 2), np.uint32)
    dist = np.empty((q.shape[0], 2), np.uint32)
    _lib().oracle_knn2_int(q.ctypes.data, r.ctypes.data, q.shape[0], r.shape[0], idx.ctypes.data,
                           dist.ctypes.data, threads)
    return idx, dist


def ratio(sorted4, thr):
    s = _f32(sorted4)
    out = np.empty((s.shape[0],), np.uint32)
    _lib().oracle_ratio(s.ctypes.data, s.shape[0], float(thr), out.ctypes.data)
    return out


def lsh_codes(x, H):
    x, H = _f32(x), _f32(H)
    out = np.empty((x.shape[0],), np.uint32)
    _lib().oracle_lsh_codes(x.ctypes.data, x.shape[0], H.ctypes.data, H.shape[0], out.ctypes.data)
    return out


def lsh_sort(codes):
    codes = np.ascontiguousarray(codes, np.uint32)
    perm = np.empty(codes.shape, np.uint32)
    _lib().oracle_lsh_sort(codes.ctypes.data, codes.shape[0], perm.ctypes.data)
    return perm


def lsh_query(q, r, codes_q, codes_r, threads=8):
    """codes_q [L, M], codes_r [L, N] -> idx, dist, ncand."""
    q, r = _f32(q), _f32(r)
    cq = np.ascontiguousarray(codes_q, np.uint32)
    cr = np.ascontiguousarray(codes_r, np.uint32)
    L = cq.shape[0]
    idx = np.empty((q.shape[0], 2), np.uint32)
    dist = np.empty((q.shape[0], 2), np.uint32)
    nc = np.empty((q.shape[0],), np.uint32)
    _lib().oracle_lsh_query(q.ctypes.data, r.ctypes.data, q.shape[0], r.shape[0], L,
                            cq.ctypes.data, cr.ctypes.data, idx.ctypes.data, dist.ctypes.data,
                            nc.ctypes.data, threads)
    return idx, dist, nc


def lsh_quantize(x, k):
    """Float-mode hash input: clamp(rint(x * 2^k), -255, 255)."""
    x = _f32(x)
    out = np.empty(x.shape, np.float32)
    _lib().oracle_lsh_quantize(x.ctypes.data, x.size, int(k), out.ctypes.data)
    return out


def lsh_scale_exp(x):
    x = _f32(x)
    return int(_lib().oracle_lsh_scale_exp(x.ctypes.data, x.size))


def lsh_query_f32(q, r, codes_q, codes_r, threads=8):
    """Float mode: codes_q [L, M], codes_r [L, N] -> idx, dist (float32, host_lenght), ncand."""
    q, r = _f32(q), _f32(r)
    cq = np.ascontiguousarray(codes_q, np.uint32)
    cr = np.ascontiguousarray(codes_r, np.uint32)
    idx = np.empty((q.shape[0], 2), np.uint32)
    dist = np.empty((q.shape[0], 2), np.float32)
    nc = np.empty((q.shape[0],), np.uint32)
    _lib().oracle_lsh_query_f32(q.ctypes.data, r.ctypes.data, q.shape[0], r.shape[0], cq.shape[0], cq.ctypes.data,
                                cr.ctypes.data, idx.ctypes.data, dist.ctypes.data, nc.ctypes.data, threads)
    return idx, dist, nc


def mlp_forward(x, W1, b1, W2, b2):
    x, W1, b1, W2, b2 = map(_f32, (x, W1, b1, W2, b2))
    n, H, B = x.shape[0], W1.shape[0], W2.shape[0]
    hid = np.empty((n, H), np.float64)
    out = np.empty((n, B), np.float64)
    _lib().oracle_mlp_forward(x.ctypes.data, n, W1.ctypes.data, b1.ctypes.data, H, W2.ctypes.data,
                              b2.ctypes.data, B, hid.ctypes.data, out.ctypes.data)
    return hid, out


HASH_NPARAM = 256 * 128 + 256 + 32 * 256 + 32


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def hash_forward(params, nbits, x, quant=True):
    """Learned hash forward (library parameter layout): (hid[n,256], z[n,32])."""
    params, x = _f32(params), _f32(x)
    n = x.shape[0]
    hid = np.empty((n, 256), np.float64)
    z = np.empty((n, 32), np.float64)
    _lib().oracle_hash_forward(params.ctypes.data, nbits, x.ctypes.data, n, 1 if quant else 0,
                               hid.ctypes.data, z.ctypes.data)
    return hid, z


def hash_loss(s, z, B, nbits, alpha, beta, gamma):
    s, z = _f64(s), _f64(z)
    return _lib().oracle_hash_loss(s.ctypes.data, z.ctypes.data, B, nbits, alpha, beta, gamma)


def hash_loss_grad(z, B, nbits, alpha, beta, gamma):
    """(loss, dL/ds at s = sign z, dL/dz with the straight-through gradient)."""
    z = _f64(z)
    dls = np.empty_like(z)
    dz = np.empty_like(z)
    L = _lib().oracle_hash_loss_grad(z.ctypes.data, B, nbits, alpha, beta, gamma, dls.ctypes.data,
                                     dz.ctypes.data)
    return L, dls, dz


def hash_backward(params, nbits, x, hid, dz, quant=True):
    params, x = _f32(params), _f32(x)
    g = np.empty(HASH_NPARAM, np.float64)
    _lib().oracle_hash_backward(params.ctypes.data, nbits, x.ctypes.data, x.shape[0], 1 if quant else 0,
                                _f64(hid).ctypes.data, _f64(dz).ctypes.data, g.ctypes.data)
    return g
