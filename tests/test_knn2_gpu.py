"""GPU parity of the exact 2-NN path against the oracle (host_brute semantics,
src/knn_brute.cu:801-885): bit-exact indices and distances for integer data,
lowest index on ties, through the C ABI."""
import os
import sys

import numpy as np
import pytest

import oracle
from matching_sift_amd import (BF16_DEVICE, COMPUTE_BF16, COMPUTE_F16, COMPUTE_F32, COMPUTE_I8,
                               COMPUTE_AUTO, COMPUTE_FP8, FLOAT_DEVICE, FLOAT_HOST, HALF_DEVICE, U8_DEVICE,
                               DeviceArray, BruteForceIndex, SiftError, knn2, datagen)
from matching_sift_amd import api

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
FMAX = np.float32(np.finfo(np.float32).max)
KINDS = [COMPUTE_BF16, COMPUTE_F16, COMPUTE_F32, COMPUTE_I8]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def expected(q, r):
    s = oracle.host_brute(q, r, threads=8)
    return s, oracle.ratio(s, 0.8)


def to_idx_dist(s):
    s64 = s.astype(np.float64)
    idx = np.where(s64[:, 2:4] == np.float64(FMAX), 4294967295.0, s64[:, 2:4]).astype(np.uint32)
    return idx, s[:, 0:2]


@pytest.mark.parametrize("compute", KINDS)
def test_config0_golden(handle, config0, compute):
    q, b = config0
    say("config0 golden compute", compute)
    idx, dist, srt, mat = knn2(handle, q, b, compute=compute, ratio=0.8)
    gold = np.load(os.path.join(GOLD, "c0_sorted.npy"))
    np.testing.assert_array_equal(srt, gold)
    gi, gd = to_idx_dist(gold)
    np.testing.assert_array_equal(idx, gi)
    np.testing.assert_array_equal(dist, gd)
    np.testing.assert_array_equal(mat, np.load(os.path.join(GOLD, "c0_matches_r08.npy")))


@pytest.mark.parametrize("compute", KINDS)
@pytest.mark.parametrize("n,m", [(1, 5), (2, 3), (31, 1), (33, 70), (64, 257), (100, 1000),
                                 (1000, 513), (4097, 300)])
def test_ragged_sizes_with_ties(handle, compute, n, m):
    say("ragged", compute, n, m)
    rng = np.random.default_rng(n * 1000 + m)
    r = rng.integers(0, 3, size=(n, 128)).astype(np.float32)  # many exact ties
    if n > 4:
        r[n // 2:n // 2 + 2] = r[0:2]  # duplicates
    q = rng.integers(0, 3, size=(m, 128)).astype(np.float32)
    q[: min(m, n)] = r[: min(m, n)]
    s, mt = expected(q, r)
    idx, dist, srt, mat = knn2(handle, q, r, compute=compute, ratio=0.8)
    np.testing.assert_array_equal(srt, s)
    np.testing.assert_array_equal(mat, mt)


@pytest.mark.parametrize("compute", KINDS)
def test_extremes(handle, compute):
    """Largest distances (0 vs 255 everywhere) and all-zero rows."""
    say("extremes", compute)
    r = np.zeros((40, 128), np.float32)
    r[1::3] = 255.0
    r[5, 7] = 1.0
    q = np.stack([np.zeros(128), np.full(128, 255.0), np.eye(1, 128, 7)[0] * 255.0]).astype(np.float32)
    s, mt = expected(q, r)
    idx, dist, srt, mat = knn2(handle, q, r, compute=compute, ratio=0.8)
    np.testing.assert_array_equal(srt, s)
    np.testing.assert_array_equal(mat, mt)


def test_empty_inputs(handle):
    q = np.ones((3, 128), np.float32)
    idx, dist, srt, mat = knn2(handle, q, np.zeros((0, 128), np.float32), compute=COMPUTE_BF16)
    assert (idx == 0xFFFFFFFF).all() and (srt == FMAX).all() and (mat == 0xFFFFFFFF).all()
    out = knn2(handle, np.zeros((0, 128), np.float32), q, compute=COMPUTE_BF16)
    assert out[0].shape == (0, 2)


@pytest.mark.parametrize("type_", [FLOAT_DEVICE, HALF_DEVICE, BF16_DEVICE, U8_DEVICE])
def test_input_types(handle, config0, type_):
    q, b = config0
    say("input type", type_)
    gold = np.load(os.path.join(GOLD, "c0_sorted.npy"))
    if type_ == HALF_DEVICE:
        qq, bb = q.astype(np.float16), b.astype(np.float16)
    elif type_ == BF16_DEVICE:
        qq = (q.view(np.uint32) >> 16).astype(np.uint16)
        bb = (b.view(np.uint32) >> 16).astype(np.uint16)
    elif type_ == U8_DEVICE:
        qq, bb = q.astype(np.uint8), b.astype(np.uint8)
    else:
        qq, bb = q, b
    idx, dist, srt, mat = knn2(handle, qq, bb, compute=COMPUTE_AUTO, type_=type_)
    np.testing.assert_array_equal(srt, gold)


def test_reference_entry_points(handle, config0):
    q, b = config0
    gold = np.load(os.path.join(GOLD, "c0_sorted.npy"))
    gmat = np.load(os.path.join(GOLD, "c0_matches_r08.npy"))
    qd, bd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(b)
    say("device_brute")
    out = DeviceArray.empty((q.shape[0], 4), np.float32)
    api.device_brute(qd, bd, q.shape[0], b.shape[0], out)
    np.testing.assert_array_equal(out.numpy(), gold)
    say("cublas_2nn_f")
    out.fill(0)
    api.cublas_2nn_f(qd, bd, q.shape[0], b.shape[0], out, handle)
    np.testing.assert_array_equal(out.numpy(), gold)
    say("cublas_2nn_sift")
    m = DeviceArray.empty((q.shape[0],), np.uint32)
    api.cublas_2nn_sift(qd, bd, FLOAT_DEVICE, q.shape[0], b.shape[0], m, 0.8, handle, 1,
                        2_000_000 * 383, 17)
    np.testing.assert_array_equal(m.numpy(), gmat)
    say("cublas_2nn_sift half")
    qh, bh = DeviceArray.from_numpy(q.astype(np.float16)), DeviceArray.from_numpy(b.astype(np.float16))
    m.fill(0)
    api.cublas_2nn_sift(qh, bh, HALF_DEVICE, q.shape[0], b.shape[0], m, 0.8, handle, 1, 0, 17)
    np.testing.assert_array_equal(m.numpy(), gmat)


def test_brute_force_index_reuse(handle, config0):
    q, b = config0
    gold = np.load(os.path.join(GOLD, "c0_sorted.npy"))
    gi, gd = to_idx_dist(gold)
    ix = BruteForceIndex(handle, b, compute=COMPUTE_BF16)
    for part in (slice(0, 1000), slice(0, 37), slice(500, 1000)):
        idx, dist = ix.query(q[part])
        np.testing.assert_array_equal(idx, gi[part])
        np.testing.assert_array_equal(dist, gd[part])
    ix.close()


def test_i8_rejects_non_integer(handle):
    q = np.full((2, 128), 0.5, np.float32)
    with pytest.raises(SiftError) as e:
        knn2(handle, q, q, compute=COMPUTE_I8)
    assert e.value.code == -5


def test_deterministic_repeat(handle, config0):
    q, b = config0
    a = knn2(handle, q, b, compute=COMPUTE_BF16)
    c = knn2(handle, q, b, compute=COMPUTE_BF16)
    for x, y in zip(a, c):
        np.testing.assert_array_equal(x, y)


def test_unit_norm_float_data_f32(handle):
    """Reference test data (unit-norm floats, src/helper.cu:40-58): FP32 MFMA
    ranks by the GEMM form, so near-ties may swap; distances within 1e-5
    relative + 1e-6 absolute of host_lenght; rank agreement >= 99%."""
    r = datagen.unit(3000, seed=5)
    q = datagen.unit(300, seed=6)
    s = oracle.host_brute(q, r, threads=8)
    idx, dist, srt, mat = knn2(handle, q, r, compute=COMPUTE_F32)
    agree = (srt[:, 2] == s[:, 2]).mean()
    assert agree >= 0.99
    ok = srt[:, 2] == s[:, 2]
    np.testing.assert_allclose(srt[ok, 0], s[ok, 0], rtol=1e-5, atol=1e-6)


SHAPES = {COMPUTE_BF16: ["2,4,2", "2,4,4", "2,8,2", "4,4,2", "1,4,2", "1,8,2"],
          COMPUTE_F16: ["2,8,2", "2,4,2"], COMPUTE_F32: ["2,4,1"],
          COMPUTE_I8: ["4,4,2", "2,4,2", "2,8,2"]}


@pytest.mark.parametrize("compute,shape", [(c, s) for c in SHAPES for s in SHAPES[c]])
def test_seeded_bound_all_shapes(handle, compute, shape, monkeypatch):
    """Bases of at least 12,288 rows take the seeded chain (the first launch
    covers 4,096 rows and bounds the rest); exercise it with realistic data and
    with a 3-letter alphabet where many rows tie exactly with the bound."""
    monkeypatch.setenv("SIFT_SHAPE", shape)
    say("seeded", compute, shape)
    b = datagen.base(40_000, seed=71)
    q = datagen.queries(b, 300, seed=72)
    s, mt = expected(q, b)
    idx, dist, srt, mat = knn2(handle, q, b, compute=compute, ratio=0.8)
    np.testing.assert_array_equal(srt, s)
    np.testing.assert_array_equal(mat, mt)
    rng = np.random.default_rng(5)
    r = rng.integers(0, 3, size=(20_000, 128)).astype(np.float32)
    r[17_000:17_100] = r[100:200]  # duplicates of sample rows far beyond the sample
    qq = rng.integers(0, 3, size=(150, 128)).astype(np.float32)
    qq[:50] = r[100:150]
    s, mt = expected(qq, r)
    idx, dist, srt, mat = knn2(handle, qq, r, compute=compute, ratio=0.8)
    np.testing.assert_array_equal(srt, s)


@pytest.mark.parametrize("compute,shape", [(COMPUTE_BF16, "2,8,2"), (COMPUTE_BF16, "2,4,2"),
                                           (COMPUTE_F16, "2,4,2"), (COMPUTE_I8, "2,4,2"),
                                           (COMPUTE_F32, "2,4,1")])
@pytest.mark.parametrize("levels", ["0", "3,17,200", "16,64,256,1000", "128,1024"])
def test_seed_chain_levels(handle, compute, shape, levels, monkeypatch):
    """The base is scanned as a chain of launches over consecutive tile
    ranges, each starting from the bound of the ranges before it, and every
    launch's partials join the final merge.  Any chain (none, unaligned range
    ends, many short ranges) gives the oracle's result exactly, also when
    duplicates of early rows sit in later ranges and tie with the bound."""
    monkeypatch.setenv("SIFT_SHAPE", shape)
    monkeypatch.setenv("SIFT_SEED", levels)
    say("seed chain", compute, shape, levels)
    b = datagen.base(40_000, seed=91)
    q = datagen.queries(b, 260, seed=92)
    s, mt = expected(q, b)
    idx, dist, srt, mat = knn2(handle, q, b, compute=compute, ratio=0.8)
    np.testing.assert_array_equal(srt, s)
    np.testing.assert_array_equal(mat, mt)
    rng = np.random.default_rng(9)
    r = rng.integers(0, 3, size=(36_000, 128)).astype(np.float32)
    for lo in (90, 600, 6_000, 31_000, 33_000):  # copies of rows of the first range in every later one
        r[lo:lo + 40] = r[10:50]
    qq = rng.integers(0, 3, size=(130, 128)).astype(np.float32)
    qq[:40] = r[10:50]
    s, mt = expected(qq, r)
    idx, dist, srt, mat = knn2(handle, qq, r, compute=compute, ratio=0.8)
    np.testing.assert_array_equal(srt, s)
    np.testing.assert_array_equal(mat, mt)


@pytest.mark.parametrize("compute", [COMPUTE_BF16, COMPUTE_I8, COMPUTE_F32])
@pytest.mark.parametrize("guide", ["", "0", "1.5,16,768", "1,16,16", "4,2,8", "0.5,400,4000"])
def test_guided_slab_schedules(handle, compute, guide, monkeypatch):
    """The launches after the first split their tile range into one region
    per XCD and each region into chunks that shrink towards its end.  Every
    schedule (the default, uniform slabs, many equal small chunks, tiny
    chunks that hit the chunk cap, one chunk per region) covers every tile
    once and gives the oracle's result exactly, with row copies at chunk and
    region edges tying across launches."""
    if guide:
        monkeypatch.setenv("SIFT_GUIDE", guide)
    else:
        monkeypatch.delenv("SIFT_GUIDE", raising=False)
    say("guided slabs", compute, repr(guide))
    rng = np.random.default_rng(21)
    b = datagen.base(70_000, seed=95)
    q = datagen.queries(b, 300, seed=96)
    for lo in (4_000, 8_000, 33_000, 35_040, 52_480, 69_950):
        b[lo:lo + 30] = b[100:130]
    q[:30] = b[100:130]
    q[30:60] = b[rng.integers(0, 70_000, 30)]
    s, mt = expected(q, b)
    idx, dist, srt, mat = knn2(handle, q, b, compute=compute, ratio=0.8)
    np.testing.assert_array_equal(srt, s)
    np.testing.assert_array_equal(mat, mt)


def test_full_size_default_chain(handle):
    """The benchmark's base (1,000,000 rows: a chain of four launches over
    128, 896, 7,168 and 23,058 tiles) against the oracle on a query sample
    that includes exact copies of base rows in every range."""
    say("full size 1M")
    b = datagen.base(1_000_000)
    q = datagen.queries(b, 96, seed=93)
    q[:4] = b[[5, 9_000, 100_000, 999_999]]
    s, mt = expected(q, b)
    idx, dist, srt, mat = knn2(handle, q, b, compute=COMPUTE_BF16, ratio=0.8)
    np.testing.assert_array_equal(srt, s)
    np.testing.assert_array_equal(mat, mt)


@pytest.mark.parametrize("shape", ["2,4,2", "2,8,2", "4,4,2"])
def test_fp8_exact_domain(handle, shape, monkeypatch):
    """FP8 coarse pass + exact re-rank.  With every value in [0,15] each
    per-row-scaled e4m3 operand is exact, so the coarse keys equal the exact
    keys and the result is the oracle's, bit for bit, through the seeded chain,
    with duplicates and exact ties; also through a brute-force index."""
    monkeypatch.setenv("SIFT_SHAPE", shape)
    rng = np.random.default_rng(21)
    r = rng.integers(0, 16, size=(40_000, 128)).astype(np.float32)
    r[30_000:30_050] = r[100:150]
    q = rng.integers(0, 16, size=(300, 128)).astype(np.float32)
    q[:50] = r[100:150]
    say("fp8 exact domain", shape)
    s, mt = expected(q, r)
    idx, dist, srt, mat = knn2(handle, q, r, compute=COMPUTE_FP8, ratio=0.8)
    np.testing.assert_array_equal(srt, s)
    np.testing.assert_array_equal(mat, mt)
    ix = BruteForceIndex(handle, r, compute=COMPUTE_FP8)
    gi, gd = to_idx_dist(s)
    ii, dd = ix.query(q)
    np.testing.assert_array_equal(ii, gi)
    np.testing.assert_array_equal(dd, gd)
    ix.close()


def test_fp8_sift_recall(handle):
    """On SIFT-like data the FP8 pass is approximate: the first neighbour is
    found for at least 99% of the queries, the ratio-test output agrees for
    at least 99%, and every returned distance is the exact squared L2 of the
    returned row (the re-rank is exact)."""
    b = datagen.base(40_000, seed=95)
    q = datagen.queries(b, 500, seed=96)
    say("fp8 sift recall")
    s, mt = expected(q, b)
    idx, dist, srt, mat = knn2(handle, q, b, compute=COMPUTE_FP8, ratio=0.8)
    gi, _ = to_idx_dist(s)
    nn1 = float((idx[:, 0] == gi[:, 0]).mean())
    agree = float((mat == mt).mean())
    say("fp8 recall nn1 %.4f, 2-nn %.4f, match %.4f" % (nn1, float((idx == gi).all(1).mean()), agree))
    assert nn1 >= 0.99 and agree >= 0.99
    assert (idx != 0xFFFFFFFF).all()
    qi, bi = q.astype(np.int64), b.astype(np.int64)
    for j in (0, 1):
        d = ((qi - bi[idx[:, j]]) ** 2).sum(1)
        np.testing.assert_array_equal(dist[:, j], d.astype(np.float32))
    assert (dist[:, 0] <= dist[:, 1]).all()


@pytest.mark.parametrize("shape", ["4,4,2", "2,4,2", "2,8,2"])
def test_fp8_deterministic(handle, shape, monkeypatch):
    """The FP8 candidates do not depend on block timing: repeated calls on a
    base large enough for the seeded chain and several rounds of blocks give
    identical outputs."""
    monkeypatch.setenv("SIFT_SHAPE", shape)
    b = datagen.base(300_000, seed=97)
    q = datagen.queries(b, 2000, seed=98)
    say("fp8 determinism", shape)
    first = knn2(handle, q, b, compute=COMPUTE_FP8, ratio=0.8)
    for _ in range(3):
        again = knn2(handle, q, b, compute=COMPUTE_FP8, ratio=0.8)
        for x, y in zip(first, again):
            np.testing.assert_array_equal(x, y)


def test_fp8_rejects_non_integer(handle):
    q = np.full((2, 128), 0.5, np.float32)
    with pytest.raises(SiftError) as e:
        knn2(handle, q, q, compute=COMPUTE_FP8)
    assert e.value.code == -5


def test_random_sizes_every_exact_kind(handle):
    """Seeded random (N, M) in the ragged and tile-boundary sizes, SIFT-like or
    tie-heavy (values 0/85/170/255) data: BF16, FP16, I8, FP32 and AUTO all
    equal host_brute exactly."""
    rng = np.random.default_rng(2024)
    for it in range(14):
        N = int(rng.choice([1, 2, 31, 32, 33, 95, 1000, 4097, 20000, 70000]))
        M = int(rng.choice([1, 2, 63, 64, 65, 257, 1000]))
        if it % 2:
            r = rng.integers(0, 4, size=(N, 128)).astype(np.float32) * 85
            q = rng.integers(0, 4, size=(M, 128)).astype(np.float32) * 85
        else:
            r = datagen.base(N, seed=100 + it)
            q = datagen.queries(r, M, seed=200 + it)
        want = oracle.host_brute(q, r, threads=8)
        for kind in (COMPUTE_BF16, COMPUTE_F16, COMPUTE_I8, COMPUTE_F32, COMPUTE_AUTO):
            say("random sizes", it, N, M, kind)
            got = knn2(handle, q, r, compute=kind)[2]
            np.testing.assert_array_equal(got, want)
