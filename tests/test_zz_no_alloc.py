"""sift_bf_index_query_noalloc (no device allocation; scratch in one
device array per GPU) against sift_bf_index_query, through the C ABI: the
same outputs bit for bit, on two calls in a row with the second smaller, for
integer and float data, blocking and not; oversize and unsupported requests
are refused."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle
from matching_sift_amd import (COMPUTE_BF16, COMPUTE_F32, COMPUTE_FP8, COMPUTE_I8, FLOAT_DEVICE, FLOAT_HOST,
                               BruteForceIndex, DeviceArray, Handle, SiftError, datagen, lib)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def outputs(m):
    return (DeviceArray.empty((m, 2), np.uint32).fill(0), DeviceArray.empty((m, 2), np.float32).fill(0),
            DeviceArray.empty((m, 4), np.float32).fill(0), DeviceArray.empty((m,), np.uint32).fill(0))


def query(fn, h, ix, qd, m, blocking=1):
    o = outputs(m)
    rc = fn(h.h, ix.ix, qd.ptr, FLOAT_DEVICE, m, o[0].ptr, o[1].ptr, o[2].ptr, o[3].ptr, C.c_float(0.8), blocking)
    if rc != 0:
        raise SiftError(rc, lib().sift_last_error(h.h).decode())
    if not blocking:
        h.sync()
    return [x.numpy() for x in o]


@pytest.fixture(scope="module")
def h():
    hd = Handle(seed=9)
    yield hd
    hd.close()


@pytest.mark.parametrize("compute,n", [(COMPUTE_BF16, 20_000), (COMPUTE_I8, 300_000), (COMPUTE_BF16, 300_000)])
def test_noalloc_equals_the_allocating_query(h, compute, n):
    b = datagen.base(n, seed=701)
    q = datagen.queries(b, 4000, seed=702)
    ix = BruteForceIndex(h, b, compute=compute)
    want_all = oracle.host_brute(q[:300], b, threads=8)
    for m in (4000, 300):  # the second call smaller
        say("noalloc", compute, n, m)
        qd = DeviceArray.from_numpy(q[:m])
        ref = query(lib().sift_bf_index_query, h, ix, qd, m)
        got = query(lib().sift_bf_index_query_noalloc, h, ix, qd, m)
        for x, y in zip(ref, got):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(got[2][:300], want_all)
    ix.close()


def test_noalloc_float_data_and_async(h):
    """An F32 index over unit-norm rows (coarse keys, exact re-rank from the
    index's tiles), non-blocking calls synchronised afterwards."""
    r = datagen.unit(30_000, seed=711)
    rng = np.random.default_rng(712)
    q = (r[rng.integers(0, 30_000, 500)] + rng.normal(0, 1e-3, (500, 128))).astype(np.float32)
    ix = BruteForceIndex(h, r, compute=COMPUTE_F32)
    for m, blocking in ((500, 0), (120, 1), (64, 0)):
        say("noalloc float", m, blocking)
        qd = DeviceArray.from_numpy(q[:m])
        ref = query(lib().sift_bf_index_query, h, ix, qd, m)
        got = query(lib().sift_bf_index_query_noalloc, h, ix, qd, m, blocking)
        for x, y in zip(ref, got):
            np.testing.assert_array_equal(x, y)
        np.testing.assert_array_equal(got[2], oracle.host_brute(q[:m], r, threads=8))
    ix.close()


def test_noalloc_refuses_what_it_cannot_hold(h):
    b = datagen.base(300_000, seed=721)
    bd = DeviceArray.from_numpy(b)
    ix = BruteForceIndex(h, bd, compute=COMPUTE_BF16)
    o = outputs(300_000)
    say("noalloc oversize")
    rc = lib().sift_bf_index_query_noalloc(h.h, ix.ix, bd.ptr, FLOAT_DEVICE, 300_000, o[0].ptr, o[1].ptr, None,
                                           None, C.c_float(0.8), 1)
    assert rc == -1 and b"maximum" in lib().sift_last_error(h.h)
    rc = lib().sift_bf_index_query_noalloc(h.h, ix.ix, bd.ptr, FLOAT_HOST, 10, o[0].ptr, None, None, None,
                                           C.c_float(0.8), 1)
    assert rc == -1
    ix.close()
    small = DeviceArray.from_numpy(b[:2000])
    ix = BruteForceIndex(h, small, compute=COMPUTE_FP8)
    rc = lib().sift_bf_index_query_noalloc(h.h, ix.ix, small.ptr, FLOAT_DEVICE, 10, o[0].ptr, None, None, None,
                                           C.c_float(0.8), 1)
    assert rc == -4
    ix.close()
    # the handle still works after the refusals
    ix = BruteForceIndex(h, small, compute=COMPUTE_I8)
    got = query(lib().sift_bf_index_query_noalloc, h, ix, DeviceArray.from_numpy(b[:50]), 50)
    np.testing.assert_array_equal(got[2], oracle.host_brute(b[:50], b[:2000], threads=8))
    ix.close()
