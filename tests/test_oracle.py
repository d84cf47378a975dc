"""Pin the CPU oracle (oracle/, restating src/knn_brute.cu:801-885 and
src/lsh_h.cu:265-317) before it is trusted as the GPU checker.

The reference ships no test vectors, so each oracle function is checked
against something that does not share its code: an independent numpy brute
force (full distance matrix + lexicographic sort), hand-worked cases, closed
forms, and planted mistakes that these checks must catch.
"""
import json
import os

import numpy as np
import pytest

import oracle
from matching_sift_amd import datagen

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FMAX = np.float32(np.finfo(np.float32).max)


def numpy_top2(q, r):
    """Independent: full int64 distance matrix, lexsort by (distance, index)."""
    qi, ri = q.astype(np.int64), r.astype(np.int64)
    d = (qi * qi).sum(1)[:, None] + (ri * ri).sum(1)[None, :] - 2 * qi @ ri.T
    M, N = d.shape
    idx = np.full((M, 2), 0xFFFFFFFF, np.uint64)
    dist = np.full((M, 2), 0xFFFFFFFF, np.uint64)
    for i in range(M):
        order = np.lexsort((np.arange(N), d[i]))[:2]
        for k, j in enumerate(order):
            idx[i, k] = j
            dist[i, k] = d[i, j]
    return idx.astype(np.uint32), dist.astype(np.uint32)


def sorted_to_int(s):
    """float4 rows -> (idx, dist) uint32 with FLT_MAX mapped to UINT32_MAX."""
    s64 = s.astype(np.float64)
    none = np.float64(FMAX)
    idx = np.where(s64[:, 2:4] == none, 4294967295.0, s64[:, 2:4]).astype(np.uint32)
    dist = np.where(s64[:, 0:2] == none, 4294967295.0, s64[:, 0:2]).astype(np.uint32)
    return idx, dist


def rand_int(rng, n, hi=256, dup_from=None):
    x = rng.integers(0, hi, size=(n, 128)).astype(np.float32)
    return x


@pytest.mark.parametrize("n", [1, 2, 3, 17, 100, 257])
def test_host_brute_matches_numpy_brute_force(n):
    rng = np.random.default_rng(n)
    r = rand_int(rng, n, hi=4)  # small alphabet -> many exact ties
    r[n // 2:] = r[: n - n // 2]  # duplicated rows
    q = np.concatenate([rand_int(rng, 20, hi=4), r[:5]])
    want_idx, want_dist = numpy_top2(q, r)
    got_idx, got_dist = sorted_to_int(oracle.host_brute(q, r))
    np.testing.assert_array_equal(got_idx, want_idx)
    np.testing.assert_array_equal(got_dist, want_dist)
    i2, d2 = oracle.knn2_int(q, r)
    np.testing.assert_array_equal(i2, want_idx)
    np.testing.assert_array_equal(d2, want_dist)


def test_hand_worked_cases():
    r = np.zeros((8, 128), np.float32)
    r[1, 0] = 3.0            # d = 9 from the zero query
    r[2, 1] = 3.0            # d = 9 (tie with row 1 -> row 1 wins)
    r[3, :4] = 1.0           # d = 4
    r[5, :4] = 1.0           # duplicate of row 3
    r[4] = 255.0             # far
    r[6, 0], r[6, 1] = 2.0, 2.0   # d = 8
    r[7, 5] = 2.0            # d = 4 (ties with rows 3, 5)
    q = np.zeros((3, 128), np.float32)   # query 0: all zeros -> rows 0 (d=0)...
    q[1, :4] = 1.0           # query 1: equals rows 3 and 5
    q[2] = 255.0             # query 2: equals row 4
    s = oracle.host_brute(q, r)
    # query 0: row 0 is the zero vector (d=0); next best d=4 at rows 3,5,7 -> row 3
    assert s[0].tolist() == [0.0, 4.0, 0.0, 3.0]
    # query 1: d=0 at rows 3 and 5 -> (3, 5)
    assert s[1].tolist() == [0.0, 0.0, 3.0, 5.0]
    # query 2 (all 255): row 4 exact (d=0); the others, worked by hand:
    #   row 0: 128*255^2 = 8,323,200     rows 1,2: 127*255^2 + 252^2 = 8,321,679
    #   rows 3,5: 4*254^2 + 124*255^2 = 8,321,164 (tie -> row 3)
    #   row 6: 2*253^2 + 126*255^2 = 8,321,168   row 7: 253^2 + 127*255^2 = 8,322,184
    assert s[2].tolist() == [0.0, 8321164.0, 4.0, 3.0]
    idx, dist = oracle.knn2_int(q, r)
    assert idx.tolist() == [[0, 3], [3, 5], [4, 3]]
    assert dist.tolist() == [[0, 4], [0, 0], [0, 8321164]]


def test_sizes_zero_and_one():
    q = np.ones((2, 128), np.float32)
    s = oracle.host_brute(q, np.zeros((0, 128), np.float32))
    assert (s == FMAX).all()
    s = oracle.host_brute(q, np.zeros((1, 128), np.float32))
    assert s[:, 0].tolist() == [128.0, 128.0] and (s[:, 1] == FMAX).all()
    assert s[:, 2].tolist() == [0.0, 0.0] and (s[:, 3] == FMAX).all()


def test_invariants_on_config0(config0):
    q, b = config0
    s = oracle.host_brute(q, b, threads=8)
    assert (s[:, 0] <= s[:, 1]).all()
    assert (s[:, 2] != s[:, 3]).all()
    i1, i2 = s[:, 2].astype(np.int64), s[:, 3].astype(np.int64)
    rec1 = ((q.astype(np.int64) - b[i1].astype(np.int64)) ** 2).sum(1)
    rec2 = ((q.astype(np.int64) - b[i2].astype(np.int64)) ** 2).sum(1)
    np.testing.assert_array_equal(rec1, s[:, 0].astype(np.int64))
    np.testing.assert_array_equal(rec2, s[:, 1].astype(np.int64))
    # multi-threaded baseline gives identical results
    np.testing.assert_array_equal(s, oracle.host_brute(q, b, threads=1))


def test_host_sort_equals_host_brute():
    rng = np.random.default_rng(5)
    r = rand_int(rng, 50, hi=3)
    q = rand_int(rng, 10, hi=3)
    d = ((q[:, None, :] - r[None, :, :]) ** 2).sum(-1).astype(np.float32)
    np.testing.assert_array_equal(oracle.host_sort(d), oracle.host_brute(q, r))


@pytest.mark.parametrize("mutation", [1, 2, 3, 4])
def test_planted_mistakes_are_caught(mutation):
    """Each plausible slip in the restatement fails the numpy cross-check."""
    rng = np.random.default_rng(3)
    r = rand_int(rng, 64, hi=3)
    r[40:48] = r[0:8]
    q = np.concatenate([rand_int(rng, 30, hi=3), r[:8]])
    want_idx, want_dist = numpy_top2(q, r)
    got_idx, got_dist = sorted_to_int(oracle.host_brute_mutant(q, r, mutation))
    assert not (np.array_equal(got_idx, want_idx) and np.array_equal(got_dist, want_dist))


def test_ratio_closed_form(config0):
    q, b = config0
    s = oracle.host_brute(q, b, threads=8)
    m = oracle.ratio(s, 0.8)
    a = np.sqrt(s[:, 0].astype(np.float32))
    c = np.sqrt(s[:, 1].astype(np.float32))
    want = np.where(a / c < np.float32(0.8), s[:, 2].astype(np.uint32), 0xFFFFFFFF)
    np.testing.assert_array_equal(m, want)
    # d2 = 0 -> 0/0 = NaN -> no match
    z = np.array([[0.0, 0.0, 3.0, 4.0]], np.float32)
    assert oracle.ratio(z, 0.8)[0] == 0xFFFFFFFF


def test_golden_config0(config0):
    q, b = config0
    meta = json.load(open(os.path.join(GOLD, "golden.json")))["config0"]
    assert datagen.sha256(b) == meta["sha256_base_f32"]
    assert datagen.sha256(q) == meta["sha256_query_f32"]
    s = oracle.host_brute(q, b, threads=8)
    np.testing.assert_array_equal(s, np.load(os.path.join(GOLD, "c0_sorted.npy")))
    np.testing.assert_array_equal(oracle.ratio(s, 0.8), np.load(os.path.join(GOLD, "c0_matches_r08.npy")))
    # integer restatement agrees with the float one (closed form: exact below 2^24)
    idx, dist = oracle.knn2_int(q, b, threads=8)
    gi, gd = sorted_to_int(s)
    np.testing.assert_array_equal(idx, gi)
    np.testing.assert_array_equal(dist, gd)


def test_lsh_codes_closed_form(config0):
    q, b = config0
    H = datagen.gauss_int(16 * 128, 32.0, 127.0, seed=99).reshape(16, 128)
    codes = oracle.lsh_codes(b[:256], H)
    proj = b[:256].astype(np.int64) @ H.astype(np.int64).T
    want = ((proj < 0).astype(np.uint64) << np.arange(16, dtype=np.uint64)).sum(1).astype(np.uint32)
    np.testing.assert_array_equal(codes, want)
    np.testing.assert_array_equal(codes, np.load(os.path.join(GOLD, "lsh_codes_256x16.npy")))
    # hyperplanes are integers in [-127, 127] (exact in bf16 and int8)
    assert np.abs(H).max() <= 127 and (H == np.round(H)).all()


def test_lsh_sort_is_stable_counting_order():
    rng = np.random.default_rng(9)
    codes = rng.integers(0, 7, size=1000).astype(np.uint32)
    perm = oracle.lsh_sort(codes)
    np.testing.assert_array_equal(perm, np.argsort(codes, kind="stable").astype(np.uint32))


def test_lsh_query_matches_candidate_brute_force():
    rng = np.random.default_rng(11)
    r = rand_int(rng, 300, hi=8)
    q = rand_int(rng, 40, hi=8)
    L = 3
    cr = rng.integers(0, 4, size=(L, 300)).astype(np.uint32)
    cq = rng.integers(0, 4, size=(L, 40)).astype(np.uint32)
    idx, dist, nc = oracle.lsh_query(q, r, cq, cr)
    for i in range(40):
        cand = np.nonzero((cr == cq[:, i:i + 1]).any(0))[0]
        assert nc[i] == len(cand)
        if len(cand) == 0:
            assert idx[i].tolist() == [0xFFFFFFFF] * 2
            continue
        wi, wd = numpy_top2(q[i:i + 1], r[cand])
        exp_idx = [cand[j] if j != 0xFFFFFFFF else 0xFFFFFFFF for j in wi[0]]
        assert idx[i].tolist() == exp_idx
        assert dist[i].tolist() == wd[0].tolist()


def test_mlp_forward_closed_form():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(5, 128)).astype(np.float32)
    W1 = rng.normal(size=(16, 128)).astype(np.float32)
    b1 = rng.normal(size=16).astype(np.float32)
    W2 = rng.normal(size=(4, 16)).astype(np.float32)
    b2 = rng.normal(size=4).astype(np.float32)
    hid, out = oracle.mlp_forward(x, W1, b1, W2, b2)
    h = np.maximum(x.astype(np.float64) @ W1.T.astype(np.float64) + b1, 0)
    np.testing.assert_allclose(hid, h, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(out, h @ W2.T.astype(np.float64) + b2, rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------ learned hash --

def _hash_params(rng, nbits, scale=1.0):
    p = np.zeros(oracle.HASH_NPARAM, np.float32)
    p[:256 * 128] = rng.normal(0, 1.28 * scale, 256 * 128)
    p[256 * 128:256 * 129] = rng.normal(0, 0.5, 256)
    w2 = rng.normal(0, 0.06, (32, 256)).astype(np.float32)
    w2[nbits:] = 0
    p[256 * 129:256 * 129 + 32 * 256] = w2.ravel()
    p[256 * 129 + 32 * 256:256 * 129 + 32 * 256 + nbits] = rng.normal(0, 0.1, nbits)
    return p


def test_hash_forward_matches_reference_forward():
    """quant=0 equals the independent oracle_mlp_forward on the same weights;
    quant=1 equals numpy with bfloat16 rounding of W1, W2, x and the hidden layer."""
    rng = np.random.default_rng(11)
    nbits = 20
    p = _hash_params(rng, nbits)
    x = rng.integers(0, 256, (7, 128)).astype(np.float32)
    W1 = p[:256 * 128].reshape(256, 128) * np.float32(1 / 64)  # layer 1 uses the stored weights / 64
    b1 = p[256 * 128:256 * 129]
    W2 = p[256 * 129:256 * 129 + 32 * 256].reshape(32, 256)
    b2 = p[256 * 129 + 32 * 256:]
    hid, z = oracle.hash_forward(p, nbits, x, quant=False)
    h2, o2 = oracle.mlp_forward(x, W1, b1, W2[:nbits], b2[:nbits])
    np.testing.assert_allclose(hid, h2, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(z[:, :nbits], o2, rtol=1e-12, atol=1e-9)
    assert (z[:, nbits:] == 0).all()

    def bf(a):
        u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
        u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
        return u.astype(np.uint32).view(np.float32).astype(np.float64)
    hq = bf(np.maximum(x.astype(np.float64) @ bf(W1).T + b1, 0).astype(np.float32))
    zq = hq @ bf(W2[:nbits]).T + b2[:nbits]
    hid_q, z_q = oracle.hash_forward(p, nbits, x, quant=True)
    np.testing.assert_allclose(hid_q, hq, rtol=0, atol=0)
    np.testing.assert_allclose(z_q[:, :nbits], zq, rtol=1e-12, atol=1e-9)


def test_hash_backward_finite_differences():
    """The backward pass is the gradient of sum(G * z) (G fixed): central
    differences on W1, b1, W2, b2 entries (no rounding: quant=0)."""
    rng = np.random.default_rng(12)
    nbits = 12
    p = _hash_params(rng, nbits)
    x = rng.integers(0, 256, (6, 128)).astype(np.float32)
    G = rng.normal(size=(6, 32))
    G[:, nbits:] = 0
    hid, z = oracle.hash_forward(p, nbits, x, quant=False)
    g = oracle.hash_backward(p, nbits, x, hid, G, quant=False)
    f = lambda q: float((oracle.hash_forward(q, nbits, x, quant=False)[1] * G).sum())  # noqa: E731
    picks = np.concatenate([rng.integers(0, 256 * 128, 12), rng.integers(256 * 128, 256 * 129, 6),
                            256 * 129 + rng.integers(0, nbits, 6) * 256 + rng.integers(0, 256, 6),
                            256 * 129 + 32 * 256 + rng.integers(0, nbits, 4)])
    for i in picks:
        eps = np.float32(2.0 ** -16)  # small enough that no ReLU input crosses zero
        qp, qm = p.copy(), p.copy()
        qp[i] += eps
        qm[i] -= eps
        fd = (f(qp) - f(qm)) / (float(qp[i]) - float(qm[i]))
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_hash_loss_closed_form_and_gradients():
    """hash_loss equals its numpy statement; dL/ds and the |z| -> 1 term's
    derivative match central differences; dL/dz is their straight-through
    sum; a dropped term would fail the closed form."""
    rng = np.random.default_rng(13)
    B, nbits, al, be, ga = 8, 6, 0.7, 0.9, 0.3
    z = rng.normal(0, 1.2, (2 * B, 32))
    z[:, nbits:] = 0
    s = np.where(z >= 0, 1.0, -1.0)
    zs, ss = z[:, :nbits], s[:, :nbits]
    N = 2 * B
    C_ = ss.T @ ss / N
    want = (((ss[:B] - ss[B:]) ** 2).sum() / 4 / (B * nbits) + al * (ss.mean(0) ** 2).sum() / nbits
            + be * ((C_ ** 2).sum() - (np.diag(C_) ** 2).sum()) / nbits ** 2
            + ga * ((np.abs(zs) - 1) ** 2).sum() / (N * nbits))
    assert abs(oracle.hash_loss(s, z, B, nbits, al, be, ga) - want) < 1e-12
    L, dls, dz = oracle.hash_loss_grad(z, B, nbits, al, be, ga)
    assert abs(L - want) < 1e-12
    eps = 1e-6
    for _ in range(30):
        i, k = rng.integers(0, N), rng.integers(0, nbits)
        sp, sm = s.copy(), s.copy()
        sp[i, k] += eps
        sm[i, k] -= eps
        fd = (oracle.hash_loss(sp, z, B, nbits, al, be, ga) - oracle.hash_loss(sm, z, B, nbits, al, be, ga)) / (2 * eps)
        assert abs(fd - dls[i, k]) < 1e-7, (i, k, fd, dls[i, k])
        zp, zm = z.copy(), z.copy()
        zp[i, k] += eps
        zm[i, k] -= eps
        fdz = (oracle.hash_loss(s, zp, B, nbits, al, be, ga) - oracle.hash_loss(s, zm, B, nbits, al, be, ga)) / (2 * eps)
        assert abs(dz[i, k] - dls[i, k] - fdz) < 1e-7, (i, k)  # straight-through + the gamma term
    assert (dz[:, nbits:] == 0).all() and (dls[:, nbits:] == 0).all()


def numpy_seq_dist(q, r):
    """Independent restatement of host_lenght for every (query, row): float32
    numpy arrays, dims added one at a time (one rounding per subtraction,
    product and addition; numpy never fuses them)."""
    q, r = q.astype(np.float32), r.astype(np.float32)
    d = np.zeros((q.shape[0], r.shape[0]), np.float32)
    for i in range(128):
        c = (q[:, i][:, None] - r[:, i][None, :]).astype(np.float32)
        d = (d + (c * c).astype(np.float32)).astype(np.float32)
    return d


def test_host_lenght_is_the_sequential_float32_sum():
    """oracle.host_lenght on float data equals the numpy sequential sum bit for
    bit; a pairwise (tree) float32 sum of the same squares differs for some
    pairs, so the check pins the summation order."""
    q = datagen.unit(40, seed=31)
    r = datagen.unit(60, seed=32) * np.float32(3.7)
    want = numpy_seq_dist(q, r)
    qi, ri = np.meshgrid(np.arange(40), np.arange(60), indexing="ij")
    got = oracle.host_lenght(q, r, qi.ravel(), ri.ravel()).reshape(40, 60)
    np.testing.assert_array_equal(got, want)
    sq = ((q[:, None, :] - r[None, :, :]) ** 2).astype(np.float32)
    tree = sq.sum(axis=2, dtype=np.float32)  # numpy's pairwise summation
    assert (tree != want).any()


def test_host_brute_on_float_data_matches_numpy():
    """host_brute on unit-norm floats (the reference's own test data,
    src/helper.cu:40-58) and on SIFT / 512 (src/test.cu:103) equals the
    independent numpy top-2: sequential float32 distances, lexicographic
    (distance, index) order, with near-duplicate queries (d << 1)."""
    for r, seed in ((datagen.unit(257, seed=41), 42), (datagen.base(300, seed=43) / np.float32(512), 44)):
        rng = np.random.default_rng(seed)
        q = r[rng.integers(0, r.shape[0], 30)] + rng.normal(0, 1e-3, (30, 128)).astype(np.float32)
        q = np.concatenate([q.astype(np.float32), datagen.unit(20, seed=seed)])
        d = numpy_seq_dist(q, r)
        s = oracle.host_brute(q, r)
        for i in range(q.shape[0]):
            order = np.lexsort((np.arange(r.shape[0]), d[i]))[:2]
            assert s[i, 2] == order[0] and s[i, 3] == order[1]
            assert s[i, 0] == d[i, order[0]] and s[i, 1] == d[i, order[1]]
        assert (s[:30, 0] < 0.01).all()
        got = oracle.host_lenght(q, r, np.arange(q.shape[0]), s[:, 2].astype(np.int64))
        np.testing.assert_array_equal(got, s[:, 0])


def test_lsh_scale_exp_and_quantize():
    """Closed forms of the float mode's hash input: k is the largest integer
    with max|x| 2^k <= 255 (checked on both sides of the boundary), and the
    hashed values are numpy's round-half-even of x 2^k, clipped to [-255, 255]."""
    for mx, k in ((255.0, 0), (256.0, -1), (127.5, 1), (127.6, 0), (1.0, 7), (0.5, 8), (1e-3, 17)):
        x = np.zeros((1, 128), np.float32)
        x[0, 5] = -mx
        assert oracle.lsh_scale_exp(x) == k
        assert np.float64(mx) * 2.0 ** k <= 255 < np.float64(mx) * 2.0 ** (k + 1)
    assert oracle.lsh_scale_exp(np.zeros((3, 128), np.float32)) == 0
    x = datagen.unit(50, seed=81) * np.float32(3.0)
    k = oracle.lsh_scale_exp(x)
    x[0, :4] = np.array([0.5, 1.5, 2.5, -0.5], np.float32) * np.float32(2.0 ** -k)  # exact halves once scaled
    assert oracle.lsh_scale_exp(x) == k
    want = np.clip(np.rint(x.astype(np.float64) * 2.0 ** k), -255, 255).astype(np.float32)
    got = oracle.lsh_quantize(x, k)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(got[0, :4], [0.0, 2.0, 2.0, -0.0])  # half to even
    np.testing.assert_array_equal(oracle.lsh_quantize(np.full((1, 128), 1e6, np.float32), 0), 255.0)


def test_lsh_query_f32_matches_numpy():
    """The float mode's re-rank: over the same-bucket candidates of any table,
    the two smallest host_lenght distances by (distance, index), equal to an
    independent numpy restatement, including duplicated base rows."""
    rng = np.random.default_rng(82)
    r = datagen.unit(400, seed=83)
    r[200:210] = r[0:10]  # exact duplicates: the lower index must win
    q = (r[rng.integers(0, 400, 60)] + rng.normal(0, 1e-3, (60, 128))).astype(np.float32)
    q[:5] = r[0:5]
    L = 3
    cr = rng.integers(0, 4, size=(L, 400)).astype(np.uint32)
    cr[:, 200:210] = cr[:, 0:10]
    cq = rng.integers(0, 4, size=(L, 60)).astype(np.uint32)
    cq[:, :5] = cr[:, :5]
    d = numpy_seq_dist(q, r)
    idx, dist, nc = oracle.lsh_query_f32(q, r, cq, cr)
    for i in range(60):
        cand = np.nonzero((cr == cq[:, i:i + 1]).any(0))[0]
        assert nc[i] == len(cand)
        order = cand[np.lexsort((cand, d[i, cand]))][:2]
        want = list(order) + [0xFFFFFFFF] * (2 - len(order))
        np.testing.assert_array_equal(idx[i], want)
        for k2, j in enumerate(order):
            assert dist[i, k2] == d[i, j]
    assert (idx[:5, 0] == np.arange(5)).all() and (dist[:5, 0] == 0).all()
