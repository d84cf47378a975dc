"""Python mirror of the reference's 2-NN entry points over the C ABI.

Reference: src/knn_brute.cu (cublas_2nn_sift :24, device_brute :619,
cublas_2nn_f :523) and src/lsh_h.cu (lsh_gpu :722).  Argument meaning and
order follow the reference; the cuBLAS handle becomes a Handle.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib

_TYPE_OF_DTYPE = {np.dtype(np.float32): _lib.FLOAT_DEVICE, np.dtype(np.float16): _lib.HALF_DEVICE,
                  np.dtype(np.uint8): _lib.U8_DEVICE}


class DeviceArray:
    """A device allocation made through the library (hipMalloc)."""

    def __init__(self, nbytes, shape=None, dtype=None):
        self.nbytes = int(nbytes)
        self.shape = shape
        self.dtype = np.dtype(dtype) if dtype is not None else None
        p = C.c_void_p()
        check(lib().sift_malloc(C.byref(p), max(1, self.nbytes)))
        self.ptr = p.value

    @classmethod
    def empty(cls, shape, dtype):
        dt = np.dtype(dtype)
        n = int(np.prod(shape)) * dt.itemsize
        return cls(n, tuple(shape), dt)

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.nbytes, arr.shape, arr.dtype)
        if arr.nbytes:
            check(lib().sift_memcpy_h2d(d.ptr, arr.ctypes.data, arr.nbytes))
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().sift_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes))
        return out

    def fill(self, byte):
        check(lib().sift_memset(self.ptr, int(byte), self.nbytes))
        return self

    def free(self):
        if self.ptr:
            lib().sift_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """Page-locked, device-mapped host memory (sift_host_alloc), viewed as a
    numpy array: the FLOAT_HOST input of the reference's caller
    (src/test.cu:63-64 allocates its descriptors this way)."""

    def __init__(self, shape, dtype):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().sift_host_alloc(C.byref(p), max(1, self.nbytes)))
        self.ptr = p.value
        buf = (C.c_char * max(1, self.nbytes)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=int(np.prod(self.shape))).reshape(self.shape)

    @classmethod
    def from_numpy(cls, arr):
        arr = np.ascontiguousarray(arr)
        out = cls(arr.shape, arr.dtype)
        out.array[...] = arr
        return out

    def numpy(self):
        return self.array.copy()

    def free(self):
        if self.ptr:
            self.array = None
            lib().sift_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def to_device(arr):
    return arr if isinstance(arr, DeviceArray) else DeviceArray.from_numpy(arr)


class Handle:
    """Owns one HIP stream and grow-only scratch (replaces cublasHandle_t)."""

    def __init__(self, seed=None):
        h = C.c_void_p()
        check(lib().sift_create(C.byref(h)))
        self.h = h.value
        if seed is not None:
            check(lib().sift_set_seed(self.h, int(seed)), self.h)

    def sync(self):
        check(lib().sift_synchronize(self.h), self.h)

    def profile(self, on=True):
        check(lib().sift_profile_enable(self.h, 1 if on else 0), self.h)

    def profile_reset(self):
        check(lib().sift_profile_reset(self.h), self.h)

    def profile_read(self, kernel):
        ms, n = C.c_double(), C.c_long()
        check(lib().sift_profile_read(self.h, kernel.encode(), C.byref(ms), C.byref(n)), self.h)
        return ms.value, n.value

    def close(self):
        if self.h:
            lib().sift_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _input_type(arr, type_=None):
    if type_ is not None:
        return type_
    return _TYPE_OF_DTYPE[np.dtype(arr.dtype)]


def knn2(handle, q, r, compute=_lib.COMPUTE_AUTO, ratio=0.8, type_=None, blocking=True,
         outputs=None):
    """Exact 2-NN of q[M,128] against r[N,128].

    q, r: numpy arrays (copied to the device) or DeviceArray.  Returns numpy
    (idx uint32[M,2], dist float32[M,2], sorted float32[M,4], matches uint32[M]).
    """
    qd, rd = to_device(q), to_device(r)
    M = qd.shape[0] if qd.shape else 0
    N = rd.shape[0] if rd.shape else 0
    t = _input_type(qd, type_)
    if outputs is None:
        outputs = (DeviceArray.empty((M, 2), np.uint32), DeviceArray.empty((M, 2), np.float32),
                   DeviceArray.empty((M, 4), np.float32), DeviceArray.empty((M,), np.uint32))
    idx, dist, srt, mat = outputs
    fn = lib().sift_knn2 if blocking else lib().sift_knn2_async
    check(fn(handle.h, qd.ptr, rd.ptr, t, M, N, compute, idx.ptr, dist.ptr, srt.ptr, mat.ptr,
             float(ratio)), handle.h)
    if not blocking:
        return outputs
    return idx.numpy(), dist.numpy(), srt.numpy(), mat.numpy()


class BruteForceIndex:
    """Base packed once into the MFMA operand layout, queried many times."""

    def __init__(self, handle, r, compute=_lib.COMPUTE_AUTO, type_=None):
        self.handle = handle
        rd = to_device(r)
        self._keep = rd
        ix = C.c_void_p()
        check(lib().sift_bf_index_build(handle.h, rd.ptr, _input_type(rd, type_), rd.shape[0],
                                        compute, C.byref(ix)), handle.h)
        self.ix = ix.value

    def query(self, q, ratio=0.8, type_=None, outputs=None, blocking=True):
        qd = to_device(q)
        M = qd.shape[0]
        if outputs is None:
            outputs = (DeviceArray.empty((M, 2), np.uint32), DeviceArray.empty((M, 2), np.float32),
                       None, None)
        idx, dist = outputs[0], outputs[1]
        srt = outputs[2] if len(outputs) > 2 else None
        mat = outputs[3] if len(outputs) > 3 else None
        check(lib().sift_bf_index_query(self.handle.h, self.ix, qd.ptr, _input_type(qd, type_), M,
                                        idx.ptr if idx else None, dist.ptr if dist else None,
                                        srt.ptr if srt else None, mat.ptr if mat else None, float(ratio),
                                        1 if blocking else 0),
              self.handle.h)
        if not blocking:
            return outputs
        if mat is not None:
            return idx.numpy(), dist.numpy(), mat.numpy()
        return idx.numpy(), dist.numpy()

    def close(self):
        if self.ix and self.handle.h:
            lib().sift_bf_index_destroy(self.ix)
        self.ix = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LSHIndex:
    """Random-hyperplane LSH index (sift_lsh_build / sift_lsh_query)."""

    def __init__(self, handle, r, nbits, L, hash_kind=_lib.LSH_HYPERPLANE, seed=0, type_=None):
        self.handle = handle
        rd = to_device(r)
        self._keep = rd
        self.n, self.nbits, self.L = rd.shape[0], int(nbits), int(L)
        ix = C.c_void_p()
        check(lib().sift_lsh_build(handle.h, rd.ptr, _input_type(rd, type_), self.n, self.nbits,
                                   self.L, hash_kind, int(seed), C.byref(ix)), handle.h)
        self.ix = ix.value

    @classmethod
    def learned(cls, handle, r, mlps, type_=None):
        """Index whose table t hashes with mlps[t] (HashMLP, kept alive by the index)."""
        self = cls.__new__(cls)
        self.handle = handle
        rd = to_device(r)
        self._keep = rd
        self._mlps = list(mlps)
        self.n, self.nbits, self.L = rd.shape[0], self._mlps[0].nbits, len(self._mlps)
        arr = (C.c_void_p * self.L)(*[m.m for m in self._mlps])
        ix = C.c_void_p()
        check(lib().sift_lsh_build_learned(handle.h, rd.ptr, _input_type(rd, type_), self.n, arr, self.L,
                                           C.byref(ix)), handle.h)
        self.ix = ix.value
        return self

    def hash_input(self):
        """(float_mode, k): a float-mode index hashes clamp(rint(x * 2^k), -255, 255)."""
        fm, k = C.c_int(), C.c_int()
        check(lib().sift_lsh_hash_input(self.ix, C.byref(fm), C.byref(k)), self.handle.h)
        return bool(fm.value), k.value

    def set_probes(self, probes):
        """Multi-probe: also look up the buckets with each of the `probes`
        least confident bits flipped (0 = off)."""
        check(lib().sift_lsh_set_probes(self.ix, int(probes)), self.handle.h)
        self.probes = int(probes)

    def bucket_stat(self):
        """Mean bucket size seen by a base-like point: sum of size^2 over the
        buckets of every table / (L n), the expected candidates per lookup."""
        v = C.c_double()
        check(lib().sift_lsh_bucket_stat(self.ix, C.byref(v)), self.handle.h)
        return v.value

    def probe_codes(self, q, type_=None):
        """[L][probes + 1][q] codes looked up per query (base code first)."""
        qd = to_device(q)
        P = getattr(self, "probes", 0)
        out = DeviceArray.empty((self.L, P + 1, qd.shape[0]), np.uint32)
        check(lib().sift_lsh_probe_codes(self.handle.h, self.ix, qd.ptr, _input_type(qd, type_), qd.shape[0],
                                         out.ptr), self.handle.h)
        return out.numpy()

    def save(self, path):
        """Write the index (hash, sorted tables, gather copy) to a file."""
        check(lib().sift_lsh_save(self.ix, str(path).encode()), self.handle.h)

    @classmethod
    def load(cls, handle, path):
        """An index from sift_lsh_save, ready to query (no rebuild); a learned
        index owns copies of its networks."""
        self = cls.__new__(cls)
        self.handle = handle
        self._keep = None
        ix = C.c_void_p()
        check(lib().sift_lsh_load(handle.h, str(path).encode(), C.byref(ix)), handle.h)
        self.ix = ix.value
        n, b, L = C.c_uint32(), C.c_int(), C.c_int()
        check(lib().sift_lsh_info(self.ix, C.byref(n), C.byref(b), C.byref(L)), handle.h)
        self.n, self.nbits, self.L = n.value, b.value, L.value
        return self

    def query_fallback(self, q, bf_index, min_cand=32, ratio=0.8, type_=None, outputs=None, accept_ratio=0.0):
        """Query, then exact brute force for the queries with fewer than
        min_cand candidates or (accept_ratio > 0) whose best candidate fails
        the ratio test sqrt(d1)/sqrt(d2) < accept_ratio against the next one.
        Returns (idx, dist, ncand, matches, n_fallback)."""
        qd = to_device(q)
        M = qd.shape[0]
        if outputs is None:
            outputs = (DeviceArray.empty((M, 2), np.uint32), DeviceArray.empty((M, 2), np.float32),
                       DeviceArray.empty((M,), np.uint32), DeviceArray.empty((M,), np.uint32))
        idx, dist, nc, mat = outputs
        nf = C.c_uint32()
        check(lib().sift_lsh_query_fallback(self.handle.h, self.ix, bf_index.ix, qd.ptr, _input_type(qd, type_), M,
                                            int(min_cand), float(accept_ratio), idx.ptr, dist.ptr, nc.ptr,
                                            mat.ptr, float(ratio),
                                            C.byref(nf)), self.handle.h)
        return idx.numpy(), dist.numpy(), nc.numpy(), mat.numpy(), nf.value

    def query(self, q, ratio=0.8, type_=None, outputs=None, blocking=True):
        qd = to_device(q)
        M = qd.shape[0]
        if outputs is None:
            outputs = (DeviceArray.empty((M, 2), np.uint32), DeviceArray.empty((M, 2), np.float32),
                       DeviceArray.empty((M,), np.uint32), DeviceArray.empty((M,), np.uint32))
        idx, dist, nc, mat = outputs
        ptr = lambda d: d.ptr if d is not None else None  # noqa: E731
        check(lib().sift_lsh_query(self.handle.h, self.ix, qd.ptr, _input_type(qd, type_), M,
                                   ptr(idx), ptr(dist), ptr(nc), ptr(mat), float(ratio),
                                   1 if blocking else 0), self.handle.h)
        if not blocking:
            return outputs
        return tuple(o.numpy() if o is not None else None for o in outputs)

    def codes(self, q, type_=None):
        qd = to_device(q)
        out = DeviceArray.empty((self.L, qd.shape[0]), np.uint32)
        check(lib().sift_lsh_codes(self.handle.h, self.ix, qd.ptr, _input_type(qd, type_),
                                   qd.shape[0], out.ptr), self.handle.h)
        return out.numpy()

    def hyperplanes(self):
        out = np.empty((self.L, self.nbits, 128), np.float32)
        check(lib().sift_lsh_get_hyperplanes(self.ix, out.ctypes.data))
        return out

    def buckets(self):
        keys = np.empty((self.L * self.n,), np.uint32)
        rows = np.empty((self.L * self.n,), np.uint32)
        check(lib().sift_lsh_get_buckets(self.ix, keys.ctypes.data, rows.ctypes.data))
        return keys, rows

    def close(self):
        if self.ix and self.handle.h:
            lib().sift_lsh_destroy(self.ix)
        self.ix = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MLP_NPARAM = 256 * 128 + 256 + 32 * 256 + 32


class TrainOpts(C.Structure):
    """sift_mlp_train_opts."""
    _fields_ = [("epochs", C.c_int), ("batch", C.c_int), ("lr", C.c_float), ("alpha", C.c_float),
                ("beta", C.c_float), ("gamma", C.c_float), ("max_ratio", C.c_float), ("seed", C.c_uint64)]


class HashMLP:
    """Learned hash 128 -> 256 (ReLU) -> nbits (sift_mlp_*)."""

    def __init__(self, handle, nbits, seed=0):
        self.handle, self.nbits = handle, int(nbits)
        m = C.c_void_p()
        check(lib().sift_mlp_create(handle.h, self.nbits, int(seed), C.byref(m)), handle.h)
        self.m = m.value

    def params(self):
        p = np.empty(MLP_NPARAM, np.float32)
        check(lib().sift_mlp_get_params(self.m, p.ctypes.data), self.handle.h)
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == MLP_NPARAM
        check(lib().sift_mlp_set_params(self.m, p.ctypes.data), self.handle.h)

    def codes(self, x, type_=None, with_z=False):
        xd = to_device(x)
        n = xd.shape[0]
        c = DeviceArray.empty((n,), np.uint32)
        z = DeviceArray.empty((n, 32), np.float32) if with_z else None
        check(lib().sift_mlp_codes(self.handle.h, self.m, xd.ptr, _input_type(xd, type_), n, c.ptr,
                                   z.ptr if z is not None else None), self.handle.h)
        return (c.numpy(), z.numpy()) if with_z else c.numpy()

    def grad(self, x, y, alpha=1.0, beta=1.0, gamma=0.1, type_=None):
        """One batch: (loss[2], grads[MLP_NPARAM], z[2B, 32])."""
        xd, yd = to_device(x), to_device(y)
        B = xd.shape[0]
        w = np.array([alpha, beta, gamma], np.float32)
        g = np.empty(MLP_NPARAM, np.float32)
        loss = np.empty(2, np.float32)
        z = np.empty((2 * B, 32), np.float32)
        check(lib().sift_mlp_grad(self.handle.h, self.m, xd.ptr, yd.ptr, _input_type(xd, type_), B, w.ctypes.data,
                                  g.ctypes.data, loss.ctypes.data, z.ctypes.data), self.handle.h)
        return loss, g, z

    def train(self, x, base, epochs=10, batch=4096, lr=1e-3, alpha=1.0, beta=1.0, gamma=0.1, max_ratio=0.8,
              seed=0, type_=None):
        xd, bd = to_device(x), to_device(base)
        o = TrainOpts(int(epochs), int(batch), float(lr), float(alpha), float(beta), float(gamma),
                      float(max_ratio), int(seed))
        el = np.empty(int(epochs), np.float32)
        check(lib().sift_mlp_train(self.handle.h, self.m, xd.ptr, xd.shape[0], bd.ptr, bd.shape[0],
                                   _input_type(xd, type_), C.byref(o), el.ctypes.data), self.handle.h)
        return el

    def close(self):
        if self.m and self.handle.h:  # a closed handle already took the device state with it
            lib().sift_mlp_destroy(self.m)
        self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


PAIRS_GRAPH = 1
PAIRS_AUTOTUNE = 2


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint32))


class PairsPlan:
    """Batched image pairs (sift_pairs_plan_*): CSR query/base descriptor sets,
    pairs[P,2] = (query image, base image); one persistent launch per execute,
    optionally captured in a HIP graph and autotuned on first execute."""

    def __init__(self, handle, q_off, r_off, pairs, compute=_lib.COMPUTE_BF16,
                 type_=_lib.FLOAT_DEVICE, flags=PAIRS_GRAPH | PAIRS_AUTOTUNE):
        self.handle = handle
        self.q_off, self.r_off, self.pairs = _u32(q_off), _u32(r_off), _u32(pairs).reshape(-1, 2)
        pl = C.c_void_p()
        check(lib().sift_pairs_plan_create(handle.h, type_, compute, len(self.q_off) - 1,
                                           self.q_off.ctypes.data, len(self.r_off) - 1,
                                           self.r_off.ctypes.data, self.pairs.shape[0],
                                           self.pairs.ctypes.data, int(flags), C.byref(pl)),
              handle.h)
        self.plan = pl.value
        self.total = int(self.info()[0])

    def info(self):
        o = (C.c_int64 * 9)()
        check(lib().sift_pairs_plan_info(self.plan, o))
        return list(o)

    def set_shape(self, qw, waves, tps):
        check(lib().sift_pairs_plan_set_shape(self.plan, qw, waves, tps), self.handle.h)

    def outputs(self):
        Q = self.total
        return (DeviceArray.empty((Q, 2), np.uint32), DeviceArray.empty((Q, 2), np.float32),
                DeviceArray.empty((Q,), np.uint32))

    def execute(self, q, r, outputs=None, ratio=0.8, blocking=True):
        if outputs is None:
            outputs = self.outputs()
        ptr = lambda d: d.ptr if d is not None else None  # noqa: E731
        check(lib().sift_pairs_plan_execute(self.handle.h, self.plan, q.ptr, r.ptr, ptr(outputs[0]),
                                            ptr(outputs[1]), ptr(outputs[2]), float(ratio),
                                            1 if blocking else 0), self.handle.h)
        if not blocking:
            return outputs
        return tuple(o.numpy() if o is not None else None for o in outputs)

    def close(self):
        if self.plan and self.handle.h:
            lib().sift_pairs_plan_destroy(self.plan)
        self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def match_pairs(handle, q, q_off, r, r_off, pairs, compute=_lib.COMPUTE_AUTO, ratio=0.8,
                type_=None):
    """One-shot batched pairs (sift_match_pairs).  Returns numpy (idx[Q,2],
    dist[Q,2], matches[Q]) over the pairs' queries in pair order."""
    if type_ == _lib.FLOAT_HOST:  # host float32 rows, staged by the library
        qh = np.ascontiguousarray(q, np.float32)
        rh = np.ascontiguousarray(r, np.float32)
        qp, rp, t = qh.ctypes.data, rh.ctypes.data, type_
    else:
        qd, rd = to_device(q), to_device(r)
        qp, rp, t = qd.ptr, rd.ptr, _input_type(qd, type_)
    q_off, r_off, pairs = _u32(q_off), _u32(r_off), _u32(pairs).reshape(-1, 2)
    Q = int(sum(int(q_off[a + 1]) - int(q_off[a]) for a in pairs[:, 0]))
    outs = (DeviceArray.empty((Q, 2), np.uint32), DeviceArray.empty((Q, 2), np.float32),
            DeviceArray.empty((Q,), np.uint32))
    check(lib().sift_match_pairs(handle.h, qp, rp, t, len(q_off) - 1,
                                 q_off.ctypes.data, len(r_off) - 1, r_off.ctypes.data,
                                 pairs.shape[0], pairs.ctypes.data, compute, outs[0].ptr,
                                 outs[1].ptr, outs[2].ptr, float(ratio)), handle.h)
    return tuple(o.numpy() for o in outs)


# ---------------------------------------------------- reference-named ----

def cublas_2nn_sift(q_points, r_points, type_, q_n, r_n, matches, threshold, handle,
                    stream_n=1, use=0, warps=1):
    """src/knn_brute.cu:24 — matches[i] = idx1 if sqrt(d1)/sqrt(d2) < threshold."""
    return check(lib().cublas_2nn_sift(q_points.ptr, r_points.ptr, type_, q_n, r_n, matches.ptr,
                                       float(threshold), handle.h, stream_n, use, warps), handle.h)


def cublas_2nn_sift_batch(q_points, type_, R, q_n, r_n, matches, threshold, handle, stream=None, warps=1):
    """src/knn_brute.cu:213 — q typed by type_ (float or half, device), R fp16
    on the device; Q and dist (the reference's caller scratch) are not needed."""
    return check(lib().cublas_2nn_sift_batch(q_points.ptr, None, type_, R.ptr, q_n, r_n, None, matches.ptr,
                                             float(threshold), handle.h, stream, warps), handle.h)


def cublas_2nn_brute_f(q_points, r_points, q_n, r_n, sorted_, handle, dist=None):
    """src/knn_brute.cu:567 — as cublas_2nn_f; the caller's dist buffer is unused."""
    lib().cublas_2nn_brute_f(q_points.ptr, r_points.ptr, q_n, r_n, sorted_.ptr,
                             dist.ptr if dist is not None else None, handle.h)
    return check(lib().sift_last_status(), handle.h)


def lsh_check_file(path):
    """Check an LSH index file on the host (no device needed); raises SiftError."""
    return check(lib().sift_lsh_check_file(str(path).encode()))


def device_brute(q_points, r_points, q_n, r_n, sorted_):
    """src/knn_brute.cu:619 — exact 2-NN into float4 {d1, d2, idx1, idx2}."""
    lib().device_brute(q_points.ptr, r_points.ptr, q_n, r_n, sorted_.ptr)
    rc = lib().sift_last_status()
    return check(rc)


def cublas_2nn_f(q_points, r_points, q_n, r_n, sorted_, handle):
    """src/knn_brute.cu:523."""
    lib().cublas_2nn_f(q_points.ptr, r_points.ptr, q_n, r_n, sorted_.ptr, handle.h)
    return check(lib().sift_last_status(), handle.h)


def lsh_gpu(q_points, r_points, type_, q_n, r_n, matches, threshold, handle, stream_n, l,
            lsh_type, nbits):
    """src/lsh_h.cu:722 — build l tables of nbits bits, query, ratio-test."""
    return check(lib().lsh_gpu(q_points.ptr, r_points.ptr, type_, q_n, r_n, matches.ptr,
                               float(threshold), handle.h, stream_n, l, lsh_type, nbits), handle.h)


# --------------------------------------------------------- SIFT files ----
VECS_KIND = {"f": 0, "i": 1, "b": 2}  # .fvecs, .ivecs, .bvecs


def _vecs_kind(path, kind):
    if kind is None:
        kind = {".fvecs": "f", ".ivecs": "i", ".bvecs": "b"}.get(path[-6:], None)
        if kind is None:
            raise ValueError("cannot tell the element kind of %s (give kind='f', 'i' or 'b')" % path)
    return VECS_KIND[kind]


def vecs_info(path, kind=None):
    """(records, dimension) of a .fvecs / .ivecs / .bvecs file."""
    n, d = C.c_uint32(), C.c_uint32()
    check(lib().sift_vecs_info(path.encode(), _vecs_kind(path, kind), C.byref(n), C.byref(d)), None)
    return n.value, d.value


def read_vecs(path, n=None, first=0, cols=None, scale=1.0, kind=None):
    """Records [first, first + n) of a SIFT vector file as a numpy array:
    float32 for .fvecs / .bvecs (times scale), int32 for .ivecs; cols keeps
    the first cols elements of each record (src/test.cu:76-155)."""
    k = _vecs_kind(path, kind)
    total, d = vecs_info(path, kind)
    n = total - first if n is None else n
    cols = d if cols is None else cols
    out = np.empty((n, cols), np.int32 if k == 1 else np.float32)
    check(lib().sift_vecs_read(path.encode(), k, first, n, cols, float(scale), out.ctypes.data), None)
    return out


def recall_count(found, gt, depths=(1, 10, 20, 30, 40, 50, 60, 70, 80, 90, 100)):
    """Per depth R: how many queries have their returned match among the first
    R ground-truth ids (the reference's recall loop, src/test.cu:224-301)."""
    found = np.ascontiguousarray(found, np.uint32)
    gt = np.ascontiguousarray(gt, np.int32)
    dep = np.ascontiguousarray(depths, np.uint32)
    hits = np.zeros(len(dep), np.uint32)
    if gt.ndim != 2 or gt.shape[0] != found.shape[0]:
        raise ValueError("gt must be [queries][k]")
    check(lib().sift_recall_count(found.ctypes.data, found.shape[0], gt.ctypes.data, gt.shape[1], dep.ctypes.data,
                                  len(dep), hits.ctypes.data), None)
    return dict(zip(dep.tolist(), hits.tolist()))
