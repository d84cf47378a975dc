"""LSH and learned-hash indexes at the full size of the 1M-row workload
(base 1,000,000 x queries 10,000, generated as the 10k x 1k case): the
oracle's restatement of src/lsh_h.cu (codes :265-317, stable bucket order
:587-591, exact re-rank of same-bucket candidates) on every base row's codes
and bucket order, every query's candidate count, and the re-ranked answers of
the first 1,000 queries; for all 10,000 queries the size-independent
properties (each returned row shares a bucket with the query, its distance
is the exact one, the pair is ordered)."""
import sys

import numpy as np
import pytest

import oracle
from matching_sift_amd import BruteForceIndex, COMPUTE_BF16, DeviceArray, HashMLP, LSHIndex, datagen, knn2

pytestmark = pytest.mark.gpu
FMAX = np.float32(np.finfo(np.float32).max)
NONE = 0xFFFFFFFF
N_ORACLE = 1_000  # queries re-ranked by the oracle (each scans all 1M rows)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def data():
    say("full-size data: 1M base, 10k queries")
    b = datagen.base(1_000_000)
    q = datagen.queries(b, 10_000)
    return q, b, DeviceArray.from_numpy(b)


def bucket_counts(cr, cq):
    """Rows of table codes cr sharing each query code in cq (one table)."""
    u, cnt = np.unique(cr, return_counts=True)
    pos = np.searchsorted(u, cq)
    pos_c = np.minimum(pos, len(u) - 1)
    return np.where(u[pos_c] == cq, cnt[pos_c], 0)


def check_buckets(ix, cr, nbits):
    """Sorted keys and rows of every table: the oracle's stable order."""
    n, L = cr.shape[1], cr.shape[0]
    keys, rows = ix.buckets()
    for t in range(L):
        perm = oracle.lsh_sort(cr[t])
        np.testing.assert_array_equal(rows[t * n:(t + 1) * n], perm)
        want = cr[t][perm].astype(np.uint64)
        if nbits < 32:
            want |= np.uint64(t) << np.uint64(nbits)
        np.testing.assert_array_equal(keys[t * n:(t + 1) * n], want.astype(np.uint32))


def check_answers(q, b, cq, cr, idx, dist, nc):
    """Candidate counts of every query, the oracle's answers for the first
    N_ORACLE, and for all: returned rows are candidates at their exact
    distance, in order."""
    L = cq.shape[0]
    want_nc = sum(bucket_counts(cr[t], cq[t]) for t in range(L))  # bucket sizes summed over tables
    np.testing.assert_array_equal(nc, want_nc.astype(np.uint32))
    say("oracle re-rank of", N_ORACLE, "queries")
    oi, od, _ = oracle.lsh_query(q[:N_ORACLE], b, cq[:, :N_ORACLE], cr, threads=16)
    np.testing.assert_array_equal(idx[:N_ORACLE], oi)
    np.testing.assert_array_equal(dist[:N_ORACLE], np.where(od == NONE, FMAX, od.astype(np.float32)))
    for k in (0, 1):
        has = idx[:, k] != NONE
        r = idx[has, k]
        shared = np.zeros(r.shape, bool)
        for t in range(L):
            shared |= cr[t][r] == cq[t][has]
        assert shared.all()
        d = ((q[has].astype(np.int64) - b[r].astype(np.int64)) ** 2).sum(1)
        np.testing.assert_array_equal(dist[has, k], d.astype(np.float32))
        np.testing.assert_array_equal(dist[~has, k], FMAX)
    both = idx[:, 1] != NONE
    assert (dist[both, 0] <= dist[both, 1]).all() and (idx[both, 0] != idx[both, 1]).all()
    # a query with a candidate gets a first row; a second only with two
    # candidate slots (a row in two tables counts twice)
    np.testing.assert_array_equal(idx[:, 0] != NONE, nc >= 1)
    assert (nc[both] >= 2).all()
    if L == 1:
        np.testing.assert_array_equal(both, nc >= 2)


@pytest.mark.parametrize("nbits,L", [(8, 1), (16, 2), (24, 1)])
def test_lsh_full_size(handle, data, nbits, L):
    """b = 8: buckets of ~24k rows (bucket-ordered re-rank); b = 16: offset
    table; b = 24: binary search over the sorted keys."""
    q, b, bd = data
    say("lsh full size", nbits, L)
    ix = LSHIndex(handle, bd, nbits, L, seed=11 + nbits)
    H = ix.hyperplanes()
    cr = np.stack([oracle.lsh_codes(b, H[t]) for t in range(L)])
    cq = np.stack([oracle.lsh_codes(q, H[t]) for t in range(L)])
    np.testing.assert_array_equal(ix.codes(q), cq)
    check_buckets(ix, cr, nbits)
    idx, dist, nc, mat = ix.query(q, ratio=0.8)
    check_answers(q, b, cq, cr, idx, dist, nc)
    s4 = np.concatenate([dist, np.where(idx == NONE, FMAX, idx.astype(np.float32))], 1)
    want_m = oracle.ratio(s4, 0.8)
    want_m[idx[:, 0] == NONE] = NONE
    np.testing.assert_array_equal(mat, want_m)
    ix.close()


def test_lsh_float_mode_full_size(handle, data):
    """The reference's SIFT1M / 512 input (src/test.cu:103) at full size: the
    float mode's codes and bucket order, and the first queries' re-rank by
    host_lenght, equal the oracle's."""
    q, b, _ = data
    bf, qf = b / np.float32(512), q[:N_ORACLE] / np.float32(512)
    say("lsh float mode full size")
    ix = LSHIndex(handle, bf, 16, 2, seed=29)
    fm, k = ix.hash_input()
    assert fm and k == oracle.lsh_scale_exp(bf)
    H = ix.hyperplanes()
    bq, qq = oracle.lsh_quantize(bf, k), oracle.lsh_quantize(qf, k)
    cr = np.stack([oracle.lsh_codes(bq, H[t]) for t in range(2)])
    cq = np.stack([oracle.lsh_codes(qq, H[t]) for t in range(2)])
    np.testing.assert_array_equal(ix.codes(qf), cq)
    check_buckets(ix, cr, 16)
    idx, dist, nc, _ = ix.query(qf)
    oi, od, _ = oracle.lsh_query_f32(qf, bf, cq, cr, threads=16)
    want_nc = sum(bucket_counts(cr[t], cq[t]) for t in range(2))  # bucket sizes summed over tables
    np.testing.assert_array_equal(nc, want_nc.astype(np.uint32))
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_array_equal(dist, od)
    ix.close()


def test_learned_full_size(handle, data):
    """A learned index (two 16-bit networks) over the 1M base: buckets and
    answers equal the oracle's LSH query given the networks' codes, and the
    exact fallback answers the queries with fewer than min_cand candidates
    (the count a third of the queries stay below) as the brute-force index
    does."""
    q, b, bd = data
    say("learned index full size")
    mlps = [HashMLP(handle, 16, seed=500 + t) for t in range(2)]
    ix = LSHIndex.learned(handle, bd, mlps)
    cr = np.stack([m.codes(bd) for m in mlps])
    cq = np.stack([m.codes(q) for m in mlps])
    check_buckets(ix, cr, 16)
    idx, dist, nc, mat = ix.query(q)
    check_answers(q, b, cq, cr, idx, dist, nc)
    say("learned fallback full size")
    bf = BruteForceIndex(handle, bd, compute=COMPUTE_BF16)
    mc = int(np.sort(nc)[len(nc) // 3]) + 1
    fi, fd, fnc, fm, nf = ix.query_fallback(q, bf, min_cand=mc)
    np.testing.assert_array_equal(fnc, nc)
    low = nc < mc
    assert nf == int(low.sum()) and nf > len(q) // 3
    ei, ed, _, em = knn2(handle, q[low], bd, compute=COMPUTE_BF16)
    np.testing.assert_array_equal(fi[low], ei)
    np.testing.assert_array_equal(fd[low], ed)
    np.testing.assert_array_equal(fm[low], em)
    np.testing.assert_array_equal(fi[~low], idx[~low])
    np.testing.assert_array_equal(fd[~low], dist[~low])
    bf.close()
    ix.close()
    for m in mlps:
        m.close()
