"""GPU parity of the reference's own calls, as src/test.cu makes them, against
the oracle (host_brute, src/knn_brute.cu:801-885), through the C ABI:

- cublas_2nn_sift with type 0 (FLOAT_HOST) on page-locked, device-mapped host
  memory (src/test.cu:63-64, :210; src/knn_brute.cu:45-79), matches in device,
  pinned host and pageable host memory, at the golden ratio 0.8 and at the
  reference's own threshold 2;
- cublas_2nn_sift_batch with an fp16 base and float or half queries
  (src/knn_brute.cu:213-278);
- cublas_2nn_brute_f (src/knn_brute.cu:567);
- exact 2-NN over the 1,000,000-row base for the I8 kind (what AUTO picks for
  integer data), FP32 and FP16, with copies of base rows in every range of
  the seeded launch chain.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle
from matching_sift_amd import (COMPUTE_AUTO, COMPUTE_F16, COMPUTE_F32, COMPUTE_I8, FLOAT_DEVICE, FLOAT_HOST,
                               HALF_DEVICE, DeviceArray, PinnedArray, api, datagen, knn2, lib)

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
FMAX = np.float32(np.finfo(np.float32).max)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "c0_sorted.npy")), np.load(os.path.join(GOLD, "c0_matches_r08.npy"))


def test_cublas_2nn_sift_float_host_pinned(handle, config0, gold):
    """The reference's timed call: type 0 on cudaHostAllocMapped-style inputs
    (here sift_host_alloc), 11 arguments as at src/test.cu:210."""
    q, b = config0
    gsorted, gmat = gold
    qh, bh = PinnedArray.from_numpy(q), PinnedArray.from_numpy(b)
    M, N = q.shape[0], b.shape[0]
    for thr, want in ((0.8, gmat), (2.0, oracle.ratio(gsorted, 2.0))):
        say("cublas_2nn_sift FLOAT_HOST, device matches, threshold", thr)
        md = DeviceArray.empty((M,), np.uint32).fill(0)
        api.cublas_2nn_sift(qh, bh, FLOAT_HOST, M, N, md, thr, handle, 1, (N * 2) * 383, 17)
        np.testing.assert_array_equal(md.numpy(), want)
        say("cublas_2nn_sift FLOAT_HOST, pinned host matches, threshold", thr)
        mp = PinnedArray((M,), np.uint32)
        mp.array[:] = 0
        api.cublas_2nn_sift(qh, bh, FLOAT_HOST, M, N, mp, thr, handle, 1, (N * 2) * 383, 17)
        np.testing.assert_array_equal(mp.numpy(), want)
        say("cublas_2nn_sift FLOAT_HOST, pageable host matches, threshold", thr)
        mh = np.zeros((M,), np.uint32)
        rc = lib().cublas_2nn_sift(qh.ptr, bh.ptr, FLOAT_HOST, M, N, mh.ctypes.data, ctypes.c_float(thr),
                                   handle.h, 1, (N * 2) * 383, 17)
        assert rc == 0
        np.testing.assert_array_equal(mh, want)
    # the inputs are left as they were (the call stages copies, it does not convert in place)
    np.testing.assert_array_equal(qh.array, q)
    np.testing.assert_array_equal(bh.array, b)


def test_cublas_2nn_sift_float_host_near_duplicates(handle):
    """Type 0 with the reference's unit-norm float data (src/helper.cu:40-58)
    and near-duplicate queries: matches equal the oracle's ratio test on
    host_brute (coarse FP16 pass + exact re-rank)."""
    r = datagen.unit(5000, seed=301)
    rng = np.random.default_rng(302)
    q = (r[rng.integers(0, 5000, 150)] + rng.normal(0, 2e-3, (150, 128))).astype(np.float32)
    q = np.concatenate([q, datagen.unit(50, seed=303)])
    s = oracle.host_brute(q, r, threads=8)
    qh, rh = PinnedArray.from_numpy(q), PinnedArray.from_numpy(r)
    md = DeviceArray.empty((q.shape[0],), np.uint32)
    for thr in (0.8, 2.0):
        say("cublas_2nn_sift FLOAT_HOST unit-norm, threshold", thr)
        api.cublas_2nn_sift(qh, rh, FLOAT_HOST, q.shape[0], r.shape[0], md, thr, handle, 1, 0, 17)
        np.testing.assert_array_equal(md.numpy(), oracle.ratio(s, thr))


def test_cublas_2nn_sift_batch(handle, config0, gold):
    """fp16 base R on the device, queries float or half (src/knn_brute.cu:213):
    integer data equal the golden; unit-norm float data equal the oracle on
    the inputs as given (R as fp16 values, q as float or half)."""
    q, b = config0
    _, gmat = gold
    M, N = q.shape[0], b.shape[0]
    R = DeviceArray.from_numpy(b.astype(np.float16))
    m = DeviceArray.empty((M,), np.uint32)
    for t, qq in ((FLOAT_DEVICE, q), (HALF_DEVICE, q.astype(np.float16))):
        say("cublas_2nn_sift_batch golden, q type", t)
        m.fill(0)
        api.cublas_2nn_sift_batch(DeviceArray.from_numpy(qq), t, R, M, N, m, 0.8, handle)
        np.testing.assert_array_equal(m.numpy(), gmat)
    r = datagen.unit(3000, seed=311).astype(np.float16)
    rng = np.random.default_rng(312)
    qf = (r[rng.integers(0, 3000, 120)].astype(np.float32) + rng.normal(0, 1e-2, (120, 128))).astype(np.float32)
    qf = np.concatenate([qf, datagen.unit(40, seed=313)])
    Rd = DeviceArray.from_numpy(r)
    mm = DeviceArray.empty((qf.shape[0],), np.uint32)
    for t, qq in ((FLOAT_DEVICE, qf), (HALF_DEVICE, qf.astype(np.float16))):
        say("cublas_2nn_sift_batch unit-norm, q type", t)
        want = oracle.ratio(oracle.host_brute(qq.astype(np.float32), r.astype(np.float32), threads=8), 0.8)
        api.cublas_2nn_sift_batch(DeviceArray.from_numpy(qq), t, Rd, qf.shape[0], 3000, mm, 0.8, handle)
        np.testing.assert_array_equal(mm.numpy(), want)


def test_cublas_2nn_brute_f(handle, config0, gold):
    q, b = config0
    gsorted, _ = gold
    qd, bd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(b)
    out = DeviceArray.empty((q.shape[0], 4), np.float32).fill(0)
    say("cublas_2nn_brute_f")
    api.cublas_2nn_brute_f(qd, bd, q.shape[0], b.shape[0], out, handle)
    np.testing.assert_array_equal(out.numpy(), gsorted)
    # the reference's caller-owned distance buffer is accepted and left alone
    dist = DeviceArray.empty((4,), np.float32).fill(0x7F)
    out.fill(0)
    api.cublas_2nn_brute_f(qd, bd, q.shape[0], b.shape[0], out, handle, dist=dist)
    np.testing.assert_array_equal(out.numpy(), gsorted)
    assert (dist.numpy().view(np.uint32) == 0x7F7F7F7F).all()


@pytest.fixture(scope="module")
def full_size():
    """The 1,000,000-row base and 128 queries, 8 of them exact copies of base
    rows in each range of the default launch chain (tiles [0,128), [128,1024),
    [1024,8192), [8192,31250): rows 0-4,095, 4,096-32,767, 32,768-262,143,
    262,144-999,999), and the oracle's answer."""
    b = datagen.base(1_000_000)
    q = datagen.queries(b, 128, seed=94)
    rows = [5, 4_095, 4_096, 9_000, 32_767, 32_768, 100_000, 262_143, 262_144, 500_000, 999_999]
    q[:len(rows)] = b[rows]
    say("oracle on 128 queries x 1M rows")
    s = oracle.host_brute(q, b, threads=16)
    return q, b, s


@pytest.mark.parametrize("compute", [COMPUTE_I8, COMPUTE_AUTO, COMPUTE_F32, COMPUTE_F16])
def test_full_size_exact_kinds(handle, full_size, compute):
    q, b, s = full_size
    say("full size 1M, compute", compute)
    bd = DeviceArray.from_numpy(b)
    idx, dist, srt, mat = knn2(handle, q, bd, compute=compute, ratio=0.8)
    np.testing.assert_array_equal(srt, s)
    np.testing.assert_array_equal(mat, oracle.ratio(s, 0.8))
