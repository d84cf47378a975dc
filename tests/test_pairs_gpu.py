"""GPU parity of batched image pairs (sift_pairs_plan_*, sift_match_pairs).

Each pair must equal the oracle's host_brute on that pair alone
(src/knn_brute.cu:801-885: squared L2, lowest index on ties) with indices
relative to the pair's base image, plus the ratio test of cublas_2nn_sift
(src/knn_brute.cu:340-420, sqrt(d1)/sqrt(d2) < threshold on true L2),
bit-exact for integer data, for every compute kind and tile shape, eager and
replayed from a HIP graph."""
import sys

import numpy as np
import pytest

import oracle
from matching_sift_amd import (COMPUTE_AUTO, COMPUTE_BF16, COMPUTE_F16, COMPUTE_F32, COMPUTE_I8, COMPUTE_FP8,
                               FLOAT_DEVICE, FLOAT_HOST, DeviceArray, PairsPlan, SiftError, datagen,
                               knn2, match_pairs, PAIRS_GRAPH, PAIRS_AUTOTUNE)

pytestmark = pytest.mark.gpu
FMAX = np.float32(np.finfo(np.float32).max)
UMAX = np.uint32(0xFFFFFFFF)
KINDS = [COMPUTE_BF16, COMPUTE_F16, COMPUTE_F32, COMPUTE_I8]
SHAPES = {COMPUTE_BF16: [(2, 8, 2), (2, 4, 2), (2, 4, 4), (4, 4, 2), (1, 4, 2), (1, 8, 2)],
          COMPUTE_F16: [(2, 8, 2), (2, 4, 2)], COMPUTE_F32: [(2, 4, 1)],
          COMPUTE_I8: [(2, 4, 2), (4, 4, 2), (2, 8, 2)]}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def expected_pairs(q, q_off, r, r_off, pairs, ratio=0.8):
    """Per-pair oracle, concatenated in pair order."""
    idx, dist, mat = [], [], []
    for a, b in np.asarray(pairs).reshape(-1, 2):
        qa = q[q_off[a]:q_off[a + 1]]
        rb = r[r_off[b]:r_off[b + 1]]
        m = qa.shape[0]
        if m == 0:
            continue
        if rb.shape[0] == 0:
            idx.append(np.full((m, 2), UMAX, np.uint32))
            dist.append(np.full((m, 2), FMAX, np.float32))
            mat.append(np.full((m,), UMAX, np.uint32))
            continue
        s = oracle.host_brute(qa, rb, threads=8)
        s64 = s.astype(np.float64)
        idx.append(np.where(s64[:, 2:4] == np.float64(FMAX), 4294967295.0, s64[:, 2:4]).astype(np.uint32))
        dist.append(s[:, 0:2].copy())
        mat.append(oracle.ratio(s, ratio))
    if not idx:
        return (np.zeros((0, 2), np.uint32), np.zeros((0, 2), np.float32), np.zeros((0,), np.uint32))
    return np.concatenate(idx), np.concatenate(dist), np.concatenate(mat)


def assert_same(got, exp, what=""):
    for g, e, name in zip(got, exp, ("idx", "dist", "matches")):
        np.testing.assert_array_equal(g, e, err_msg="%s %s" % (what, name))


def ragged_batch(seed, ties=False):
    """Query and base image lists of different lengths, sizes around the tile
    and group edges (0, 1, 31, 32, 33, ...), pairs that reuse images."""
    rng = np.random.default_rng(seed)
    q_sizes = [0, 1, 31, 32, 33, 257, 513, 700, 1100, 64, 5]
    r_sizes = [1, 2, 31, 32, 33, 0, 640, 1500, 96, 4100]
    pairs = [(1, 0), (2, 1), (3, 2), (4, 3), (5, 4), (6, 6), (7, 7), (8, 8), (9, 9), (0, 3),
             (10, 5), (2, 5), (4, 9), (7, 2), (8, 7), (6, 6), (3, 8)]

    def images(sizes, base_seed):
        if ties:  # 3-letter alphabet: many exact distance ties
            parts = [rng.integers(0, 3, size=(n, 128)).astype(np.float32) for n in sizes]
        else:
            parts = [datagen.base(n, seed=base_seed + i) for i, n in enumerate(sizes)]
        off = np.zeros(len(sizes) + 1, np.uint32)
        off[1:] = np.cumsum(sizes)
        return np.concatenate(parts).astype(np.float32), off

    r, r_off = images(r_sizes, 500 + seed)
    if ties:
        q, q_off = images(q_sizes, 700 + seed)
    else:  # half the queries near rows of their pair's base image
        q, q_off = images(q_sizes, 700 + seed)
        for a, b in pairs:
            na, nb = int(q_off[a + 1] - q_off[a]), int(r_off[b + 1] - r_off[b])
            if na and nb:
                src = rng.integers(0, nb, size=na // 2)
                noisy = r[r_off[b] + src] + rng.normal(0, 8, size=(na // 2, 128))
                q[q_off[a]:q_off[a] + na // 2] = np.clip(np.rint(noisy), 0, 255)
    return q, q_off, r, r_off, np.array(pairs, np.uint32)


@pytest.mark.parametrize("compute", KINDS)
@pytest.mark.parametrize("ties", [False, True])
def test_pairs_all_shapes(handle, compute, ties):
    q, q_off, r, r_off, pairs = ragged_batch(3 if ties else 4, ties=ties)
    exp = expected_pairs(q, q_off, r, r_off, pairs)
    qd, rd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(r)
    plan = PairsPlan(handle, q_off, r_off, pairs, compute=compute, flags=0)
    assert plan.total == exp[0].shape[0]
    assert plan.info()[8] == 0
    for shape in SHAPES[compute]:
        say("pairs", compute, "ties" if ties else "natural", shape)
        plan.set_shape(*shape)
        assert tuple(plan.info()[3:6]) == shape
        assert_same(plan.execute(qd, rd), exp, "shape %s" % (shape,))
    plan.close()


def test_pairs_graph_replay(handle):
    """Replays of the captured graph, with eager work in between, each equal
    to the eager result; new output pointers re-capture."""
    q, q_off, r, r_off, pairs = ragged_batch(5)
    exp = expected_pairs(q, q_off, r, r_off, pairs)
    qd, rd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(r)
    plan = PairsPlan(handle, q_off, r_off, pairs, compute=COMPUTE_BF16,
                     flags=PAIRS_GRAPH | PAIRS_AUTOTUNE)
    outs = plan.outputs()
    for o in outs:
        o.fill(0xAB)
    say("graph first execute (autotune + capture)")
    assert_same(plan.execute(qd, rd, outputs=outs), exp, "first")
    info = plan.info()
    assert info[6] == 1, "graph not captured"
    say("autotuned shape", info[3:6])
    rb = datagen.base(3000, seed=91)
    qb = datagen.queries(rb, 200, seed=92)
    for rep in range(4):
        for o in outs:
            o.fill(0x5C)  # stale values must be overwritten by the replay
        say("graph replay", rep)
        assert_same(plan.execute(qd, rd, outputs=outs), exp, "replay %d" % rep)
        knn2(handle, qb, rb, compute=COMPUTE_BF16)  # unrelated eager work between replays
    outs2 = plan.outputs()
    say("graph re-capture for new outputs")
    assert_same(plan.execute(qd, rd, outputs=outs2), exp, "recapture")
    # ratio change re-captures too
    exp5 = expected_pairs(q, q_off, r, r_off, pairs, ratio=0.5)
    assert_same(plan.execute(qd, rd, outputs=outs2, ratio=0.5), exp5, "ratio 0.5")
    plan.close()


def test_pairs_config3_sample(handle):
    """48 pairs with the batched-pairs size distribution (LogNormal median
    2000, clipped to [200, 8000]; 40% noisy copies), autotuned + graph."""
    sizes = datagen.pair_sizes(48, seed=123)
    q, q_off, r, r_off, pairs = datagen.pairs_batch(48, seed=31000, sizes=sizes)
    say("config3 sample", q.shape, r.shape)
    exp = expected_pairs(q, q_off, r, r_off, pairs)
    qd, rd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(r)
    plan = PairsPlan(handle, q_off, r_off, pairs, compute=COMPUTE_BF16)
    outs = plan.outputs()
    for rep in range(2):
        assert_same(plan.execute(qd, rd, outputs=outs), exp, "rep %d" % rep)
    assert (exp[2] != UMAX).mean() > 0.2  # the noisy copies pass the ratio test
    plan.close()


@pytest.mark.parametrize("compute", [COMPUTE_I8, COMPUTE_BF16])
def test_pairs_config3_full(handle, compute):
    """All 512 pairs of the batched-pairs workload (1.21 M queries), autotuned
    and graph-replayed: every pair equals the brute-force kernel on that pair
    alone (itself bit-exact against the oracle), and the first 16 pairs equal
    the oracle's host_brute directly."""
    q, q_off, r, r_off, pairs = datagen.config3()
    say("config3 full", compute, q.shape, r.shape)
    qd, rd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(r)
    plan = PairsPlan(handle, q_off, r_off, pairs, compute=compute)
    outs = plan.outputs()
    got = None
    for rep in range(2):  # the first execute autotunes and captures the graph
        got = plan.execute(qd, rd, outputs=outs)
    gi, gd, gm = got
    n16 = int(q_off[16])
    assert_same((gi[:n16], gd[:n16], gm[:n16]), expected_pairs(q, q_off, r, r_off, pairs[:16]), "oracle")
    say("per-pair brute force on 512 pairs")
    for p in range(len(pairs)):
        a, b = int(q_off[p]), int(q_off[p + 1])
        ei, ed, _, em = knn2(handle, q[a:b], r[r_off[p]:r_off[p + 1]], compute=compute, ratio=0.8)
        assert_same((gi[a:b], gd[a:b], gm[a:b]), (ei, ed, em), "pair %d" % p)
    assert (gm != UMAX).mean() > 0.2
    plan.close()


@pytest.mark.parametrize("seed", [5, 9, 13])
def test_pairs_random_batches(handle, seed):
    """Random batches: 1-5 images a side with sizes at tile and group edges
    (empty images included), 1-5 pairs that may repeat, natural or
    3-letter-alphabet data; every exact kind, one-shot."""
    rng = np.random.default_rng(seed)
    q_pick = [0, 1, 5, 31, 32, 33, 100, 700, 2049]
    r_pick = [0, 1, 2, 31, 32, 33, 100, 900, 3000]
    done = 0
    while done < 12:
        qs = [int(rng.choice(q_pick)) for _ in range(int(rng.integers(1, 6)))]
        rs = [int(rng.choice(r_pick)) for _ in range(int(rng.integers(1, 6)))]
        n_p = int(rng.integers(1, 6))
        pairs = np.stack([rng.integers(0, len(qs), n_p), rng.integers(0, len(rs), n_p)], 1).astype(np.uint32)
        if sum(qs) == 0 or sum(rs) == 0 or sum(qs[a] for a, _ in pairs) == 0:
            continue
        ties = bool(rng.integers(0, 2))
        if ties:
            q = (rng.integers(0, 3, size=(sum(qs), 128)) * 85).astype(np.float32)
            r = (rng.integers(0, 3, size=(sum(rs), 128)) * 85).astype(np.float32)
        else:
            q = datagen.base(sum(qs), seed=int(rng.integers(1 << 30)))
            r = datagen.base(sum(rs), seed=int(rng.integers(1 << 30)))
        q_off = np.concatenate([[0], np.cumsum(qs)]).astype(np.uint32)
        r_off = np.concatenate([[0], np.cumsum(rs)]).astype(np.uint32)
        exp = expected_pairs(q, q_off, r, r_off, pairs)
        for c in KINDS:
            say("random batch", seed, done, "ties" if ties else "natural", qs, rs, pairs.tolist(), "kind", c)
            assert_same(match_pairs(handle, q, q_off, r, r_off, pairs, compute=c), exp,
                        "batch %d kind %d" % (done, c))
        done += 1


def test_match_pairs_oneshot(handle):
    q, q_off, r, r_off, pairs = ragged_batch(6)
    exp = expected_pairs(q, q_off, r, r_off, pairs)
    say("one-shot AUTO (device input)")
    assert_same(match_pairs(handle, q, q_off, r, r_off, pairs), exp, "auto")
    say("one-shot host input")
    assert_same(match_pairs(handle, q, q_off, r, r_off, pairs, type_=FLOAT_HOST), exp, "host")
    for c in (COMPUTE_F32, COMPUTE_I8):
        say("one-shot compute", c)
        assert_same(match_pairs(handle, q, q_off, r, r_off, pairs, compute=c), exp, "compute %d" % c)


def test_match_pairs_non_integer(handle):
    """AUTO falls back to FP32 for non-integer data (agrees with the oracle
    up to FP32 rounding of the expanded form); I8 refuses it."""
    rng = np.random.default_rng(8)
    q = rng.random((300, 128), dtype=np.float32)
    r = rng.random((900, 128), dtype=np.float32)
    q_off = np.array([0, 100, 300], np.uint32)
    r_off = np.array([0, 400, 900], np.uint32)
    pairs = np.array([[0, 0], [1, 1], [0, 1]], np.uint32)
    exp = expected_pairs(q, q_off, r, r_off, pairs)
    say("non-integer AUTO")
    idx, dist, _ = match_pairs(handle, q, q_off, r, r_off, pairs)
    assert (idx[:, 0] == exp[0][:, 0]).mean() > 0.99
    np.testing.assert_allclose(dist, exp[1], rtol=1e-4, atol=1e-3)
    for c in (COMPUTE_I8, COMPUTE_BF16, COMPUTE_F16):
        with pytest.raises(SiftError):
            match_pairs(handle, q, q_off, r, r_off, pairs, compute=c)
    # a plan records it in its info
    qd, rd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(r)
    plan = PairsPlan(handle, q_off, r_off, pairs, compute=COMPUTE_F32, flags=0)
    idx32, dist32, _ = plan.execute(qd, rd)
    assert plan.info()[8] == 1
    np.testing.assert_array_equal(idx32, idx)
    plan.close()


def test_pairs_bad_params(handle):
    q_off = np.array([0, 10, 20], np.uint32)
    r_off = np.array([0, 5], np.uint32)
    with pytest.raises(SiftError):
        PairsPlan(handle, q_off, r_off, np.array([[0, 1]], np.uint32))  # base image 1 missing
    with pytest.raises(SiftError):
        PairsPlan(handle, np.array([0, 10, 5], np.uint32), r_off, np.array([[0, 0]], np.uint32))
    with pytest.raises(SiftError):
        PairsPlan(handle, q_off, r_off, np.array([[0, 0]], np.uint32), compute=COMPUTE_AUTO)
    with pytest.raises(SiftError):
        PairsPlan(handle, q_off, r_off, np.array([[0, 0]], np.uint32), type_=FLOAT_HOST)
    plan = PairsPlan(handle, q_off, r_off, np.array([[0, 0]], np.uint32), flags=0)
    with pytest.raises(SiftError):
        plan.set_shape(3, 3, 3)
    plan.close()
    # no pairs: nothing to do
    empty = PairsPlan(handle, q_off, r_off, np.zeros((0, 2), np.uint32), type_=FLOAT_DEVICE)
    assert empty.total == 0
    empty.close()


FP8_SHAPES = [(2, 4, 2), (2, 8, 2), (4, 4, 2)]


def exact_dist(q, r, qi, ri):
    d = q[qi].astype(np.int64) - r[ri].astype(np.int64)
    return (d * d).sum(-1)


@pytest.mark.parametrize("ties", [False, True])
def test_pairs_fp8_exact_domain(handle, ties):
    """Values in [0, 15] are exact in e4m3 at any power-of-two scale and the
    dot products stay far below 2^24, so the coarse keys equal the true keys;
    the two nearest rows are then always among the 16 candidates (each is
    top-2 of its quarter and lane half) and the re-ranked result is exact."""
    q, q_off, r, r_off, pairs = ragged_batch(7 if ties else 8, ties=ties)
    q = np.minimum(q, 15.0).astype(np.float32)
    r = np.minimum(r, 15.0).astype(np.float32)
    exp = expected_pairs(q, q_off, r, r_off, pairs)
    qd, rd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(r)
    plan = PairsPlan(handle, q_off, r_off, pairs, compute=COMPUTE_FP8, flags=0)
    for shape in FP8_SHAPES:
        say("fp8 exact domain", "ties" if ties else "natural", shape)
        plan.set_shape(*shape)
        assert_same(plan.execute(qd, rd), exp, "fp8 shape %s" % (shape,))
    plan.close()


def test_pairs_fp8_sift_recall(handle):
    """SIFT-like pairs: every returned distance is the exact squared L2 of
    the returned row, and the FP8 coarse pass loses the true nearest row for
    at most 1% of queries (the loss is reported by the bench)."""
    sizes = datagen.pair_sizes(24, seed=321)
    q, q_off, r, r_off, pairs = datagen.pairs_batch(24, seed=33000, sizes=sizes)
    exp = expected_pairs(q, q_off, r, r_off, pairs)
    qd, rd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(r)
    plan = PairsPlan(handle, q_off, r_off, pairs, compute=COMPUTE_FP8)
    outs = plan.outputs()
    for rep in range(3):  # autotune + capture, then two replays
        say("fp8 sift rep", rep)
        idx, dist, mat = plan.execute(qd, rd, outputs=outs)
        qrow = np.concatenate([np.arange(q_off[a], q_off[a + 1]) for a, _ in pairs])
        boff = np.concatenate([np.full(q_off[a + 1] - q_off[a], r_off[b]) for a, b in pairs])
        for k in range(2):
            ok = idx[:, k] != UMAX
            d = exact_dist(q, r, qrow[ok], boff[ok] + idx[ok, k])
            np.testing.assert_array_equal(dist[ok, k], d.astype(np.float32))
        assert (dist[:, 0] <= dist[:, 1]).all()
        assert (idx[:, 0] != idx[:, 1]).all()
        r1 = (idx[:, 0] == exp[0][:, 0]).mean()
        say("fp8 recall@1", r1)
        assert r1 >= 0.99
        # matches follow the ratio test on the returned (exact) distances
        want = np.where((idx[:, 0] != UMAX) & (np.sqrt(dist[:, 0]) / np.sqrt(dist[:, 1]) < np.float32(0.8)),
                        idx[:, 0], UMAX)
        np.testing.assert_array_equal(mat, want)
    plan.close()
