"""GPU LSH (hyperplane projection, radix-sorted buckets, exact re-rank) against
the oracle's restatement of src/lsh_h.cu: sign codes (:265-317), stable
bucket order (:587-591), same-bucket candidates re-ranked exactly."""
import sys

import numpy as np
import pytest

import oracle
from matching_sift_amd import (FLOAT_DEVICE, FLOAT_HOST, LSH_HYPERPLANE, LSH_HYPERPLANE_PM1,
                               DeviceArray, LSHIndex, SiftError, datagen)
from matching_sift_amd import api

pytestmark = pytest.mark.gpu
FMAX = np.float32(np.finfo(np.float32).max)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_tables(x, H):
    return np.stack([oracle.lsh_codes(x, H[t]) for t in range(H.shape[0])])


def check_index(ix, q, b, ratio=0.8):
    L, nbits, n = ix.L, ix.nbits, b.shape[0]
    H = ix.hyperplanes()
    assert H.shape == (L, nbits, 128) and np.abs(H).max() <= 127 and (H == np.round(H)).all()
    cr, cq = oracle_tables(b, H), oracle_tables(q, H)
    np.testing.assert_array_equal(ix.codes(q), cq)
    np.testing.assert_array_equal(ix.codes(b[:2048]), cr[:, :2048])
    keys, rows = ix.buckets()
    for t in range(L):
        perm = oracle.lsh_sort(cr[t])
        np.testing.assert_array_equal(rows[t * n:(t + 1) * n], perm)
        want = cr[t][perm].astype(np.uint64)
        if nbits < 32:
            want |= np.uint64(t) << np.uint64(nbits)
        np.testing.assert_array_equal(keys[t * n:(t + 1) * n], want.astype(np.uint32))
    idx, dist, nc, mat = ix.query(q, ratio=ratio)
    oi, od, _ = oracle.lsh_query(q, b, cq, cr)
    np.testing.assert_array_equal(idx, oi)
    want_d = np.where(oi == 0xFFFFFFFF, FMAX, od.astype(np.float32))
    np.testing.assert_array_equal(dist, want_d)
    want_nc = sum((cr[t][None, :] == cq[t][:, None]).sum(1) for t in range(L))
    np.testing.assert_array_equal(nc, want_nc.astype(np.uint32))
    s4 = np.concatenate([want_d, np.where(oi == 0xFFFFFFFF, FMAX, oi.astype(np.float32))], 1)
    exp_m = oracle.ratio(s4, ratio)
    exp_m[oi[:, 0] == 0xFFFFFFFF] = 0xFFFFFFFF
    np.testing.assert_array_equal(mat, exp_m)
    return idx, nc


@pytest.mark.parametrize("nbits,L", [(1, 1), (8, 2), (12, 4), (16, 8), (20, 3), (24, 2), (32, 1), (8, 12), (4, 20)])
def test_lsh_matches_oracle(handle, nbits, L):
    say("lsh", nbits, L)
    b = datagen.base(12_000, seed=31)
    q = datagen.queries(b, 300, seed=32)
    ix = LSHIndex(handle, b, nbits, L, seed=7 + nbits)
    idx, nc = check_index(ix, q, b)
    ix.close()


def test_lsh_pm1_hyperplanes_and_small_bases(handle):
    for n in (1, 2, 33, 100):
        say("lsh small", n)
        b = datagen.base(n, seed=40 + n)
        q = datagen.queries(b, 50, seed=41)
        ix = LSHIndex(handle, b, 4, 3, hash_kind=LSH_HYPERPLANE_PM1, seed=3)
        H = ix.hyperplanes()
        assert set(np.unique(H)) <= {-1.0, 1.0}
        check_index(ix, q, b)
        ix.close()


def test_lsh_recall_rises_with_tables(handle):
    """More tables -> superset of candidates -> recall can only rise."""
    b = datagen.base(20_000, seed=51)
    q = datagen.queries(b, 400, seed=52)
    exact = oracle.knn2_int(q, b, threads=8)[0][:, 0]
    rec = []
    for L in (1, 2, 4, 8):
        ix = LSHIndex(handle, b, 16, L, seed=99)  # same seed: table t identical across L
        idx = ix.query(q)[0]
        rec.append(float((idx[:, 0] == exact).mean()))
        ix.close()
    say("recall by L", rec)
    assert all(x <= y for x, y in zip(rec, rec[1:]))
    assert rec[-1] > rec[0]


def test_lsh_gpu_entry_point(handle):
    b = datagen.base(8_000, seed=61)
    q = datagen.queries(b, 200, seed=62)
    qd, bd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(b)
    m = DeviceArray.empty((200,), np.uint32)
    api.lsh_gpu(qd, bd, FLOAT_DEVICE, 200, 8000, m, 0.8, handle, 16, 4, LSH_HYPERPLANE, 12)
    ix = LSHIndex(handle, b, 12, 4, seed=1234)  # the handle's seed
    np.testing.assert_array_equal(m.numpy(), ix.query(q, ratio=0.8)[3])
    ix.close()


def check_float_index(ix, q, b, ratio=0.8, P=0):
    """Float mode: the hashed rows are oracle.lsh_quantize(x, k) with k from
    the base, codes and buckets as the integer path on those, and the re-rank
    is the oracle's by host_lenght (sequential FP32) over the rows as given."""
    L, nbits, n = ix.L, ix.nbits, b.shape[0]
    fm, k = ix.hash_input()
    assert fm and k == oracle.lsh_scale_exp(b)
    H = ix.hyperplanes()
    bq, qq = oracle.lsh_quantize(b, k), oracle.lsh_quantize(q, k)
    cr, cq = oracle_tables(bq, H), oracle_tables(qq, H)
    np.testing.assert_array_equal(ix.codes(q), cq)
    np.testing.assert_array_equal(ix.codes(b), cr)
    keys, rows = ix.buckets()
    for t in range(L):
        np.testing.assert_array_equal(rows[t * n:(t + 1) * n], oracle.lsh_sort(cr[t]))
    if P:
        ix.set_probes(P)
        pc = ix.probe_codes(q)
        for t in range(L):  # probes flip the bits of smallest |h . x| of the hashed integers
            marg = np.abs(qq.astype(np.float64) @ H[t].astype(np.float64).T)
            flips = pc[t, 1:] ^ cq[t][None]
            np.testing.assert_array_equal(np.log2(flips).astype(np.int64).T, smallest_bits(marg, P))
        cq = pc.reshape(L * (P + 1), -1)
        cr = np.repeat(cr, P + 1, axis=0)
    idx, dist, nc, mat = ix.query(q, ratio=ratio)
    oi, od, _ = oracle.lsh_query_f32(q, b, cq, cr)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_array_equal(dist, od)
    want_nc = sum((cr[t][None, :] == cq[t][:, None]).sum(1) for t in range(cq.shape[0]))  # with repeats
    np.testing.assert_array_equal(nc, want_nc.astype(np.uint32))
    s4 = np.concatenate([od, np.where(oi == 0xFFFFFFFF, FMAX, oi.astype(np.float32))], 1)
    want_m = oracle.ratio(s4, ratio)
    want_m[oi[:, 0] == 0xFFFFFFFF] = 0xFFFFFFFF
    np.testing.assert_array_equal(mat, want_m)
    return idx


@pytest.mark.parametrize("nbits,L,P", [(8, 2, 0), (12, 4, 0), (16, 8, 3), (24, 2, 0), (32, 1, 2)])
def test_lsh_float_data(handle, nbits, L, P):
    """The reference's own inputs (unit-norm floats, src/helper.cu:40-58, with
    near-duplicate queries) build the float mode and equal the oracle's
    float-mode LSH exactly: codes, buckets, probes, re-rank and matches."""
    say("lsh float", nbits, L, P)
    b = datagen.unit(8000, seed=71)
    rng = np.random.default_rng(72)
    q = (b[rng.integers(0, 8000, 200)] + rng.normal(0, 2e-3, (200, 128))).astype(np.float32)
    q = np.concatenate([q, datagen.unit(100, seed=73)])
    ix = LSHIndex(handle, b, nbits, L, seed=3 + nbits)
    idx = check_float_index(ix, q, b, P=P)
    if nbits <= 12:  # large buckets: the near duplicates find their source row
        assert (idx[:200, 0] == oracle.host_brute(q[:200], b, threads=8)[:, 2]).mean() > 0.9
    ix.close()


def test_lsh_float_sift_div_512_and_half(handle):
    """SIFT / 512 (src/test.cu:103) and fp16 inputs (the reference converts
    float inputs to fp16, src/lsh_h.cu:748-780): float mode, exact against the
    oracle on the values as given."""
    b = datagen.base(6000, seed=74) / np.float32(512)
    q = datagen.queries(b * 512, 250, seed=75) / np.float32(512)
    ix = LSHIndex(handle, b, 12, 3, seed=5)
    check_float_index(ix, q, b)
    ix.close()
    bh, qh = b.astype(np.float16), q.astype(np.float16)
    ix = LSHIndex(handle, bh, 12, 3, seed=5)
    check_float_index(ix, qh.astype(np.float32), bh.astype(np.float32))
    ix.close()


def test_lsh_integer_index_refuses_float_queries(handle):
    b = datagen.base(3000, seed=76)
    ix = LSHIndex(handle, b, 12, 2, seed=1)
    assert ix.hash_input() == (False, 0)
    with pytest.raises(SiftError) as e:
        ix.query(datagen.queries(b, 20, seed=77) + np.float32(0.5))
    assert e.value.code == -5
    ix.close()


def test_lsh_gpu_float_data(handle):
    """lsh_gpu with the reference caller's unit-norm data (type 1), and with an
    integer base but non-integer queries (the float mode is then chosen for
    the base too)."""
    b = datagen.unit(6000, seed=78)
    rng = np.random.default_rng(79)
    q = (b[rng.integers(0, 6000, 150)] + rng.normal(0, 2e-3, (150, 128))).astype(np.float32)
    for base, qq in ((b, q), (datagen.base(6000, seed=80), datagen.queries(datagen.base(6000, seed=80), 150,
                                                                            seed=81) + np.float32(0.25))):
        say("lsh_gpu float")
        qd, bd = DeviceArray.from_numpy(qq), DeviceArray.from_numpy(base)
        m = DeviceArray.empty((qq.shape[0],), np.uint32)
        api.lsh_gpu(qd, bd, FLOAT_DEVICE, qq.shape[0], base.shape[0], m, 0.8, handle, 16, 4, LSH_HYPERPLANE, 12)
        k = oracle.lsh_scale_exp(base)
        ix = LSHIndex(handle, b[:64], 12, 4, seed=1234)  # the handle's seed: the same hyperplanes
        H = ix.hyperplanes()
        ix.close()
        cr = oracle_tables(oracle.lsh_quantize(base, k), H)
        cq = oracle_tables(oracle.lsh_quantize(qq, k), H)
        oi, od, _ = oracle.lsh_query_f32(qq, base, cq, cr)
        s4 = np.concatenate([od, np.where(oi == 0xFFFFFFFF, FMAX, oi.astype(np.float32))], 1)
        want = oracle.ratio(s4, 0.8)
        want[oi[:, 0] == 0xFFFFFFFF] = 0xFFFFFFFF
        np.testing.assert_array_equal(m.numpy(), want)


def test_lsh_rejects_bad_params(handle):
    b = datagen.base(64, seed=2)
    for nbits, L in ((0, 1), (33, 1), (8, 0), (31, 4)):
        with pytest.raises(SiftError):
            LSHIndex(handle, b, nbits, L)


def test_lsh_save_load(handle, tmp_path):
    """A saved index loads on the handle without a rebuild and answers
    exactly as before (direct bucket offsets at b = 12, binary search at
    b = 24); damaged files are rejected."""
    b = datagen.base(12_000, seed=61)
    q = datagen.queries(b, 300, seed=62)
    for nbits, L in ((12, 3), (24, 2)):
        say("lsh save/load", nbits, L)
        ix = LSHIndex(handle, b, nbits, L, seed=5)
        ref = ix.query(q)
        planes = ix.hyperplanes()
        p = tmp_path / ("ix%d.lsh" % nbits)
        ix.save(p)
        ix.close()
        jx = LSHIndex.load(handle, p)
        assert (jx.n, jx.nbits, jx.L) == (len(b), nbits, L)
        np.testing.assert_array_equal(jx.hyperplanes(), planes)
        for x, y in zip(ref, jx.query(q)):
            np.testing.assert_array_equal(x, y)
        jx.close()
    say("lsh save/load, float mode")
    bf = datagen.unit(5000, seed=63)
    qf = bf[:100] + np.float32(1e-3)
    ix = LSHIndex(handle, bf, 12, 3, seed=6)
    ref = ix.query(qf)
    ix.save(tmp_path / "f.lsh")
    ix.close()
    jx = LSHIndex.load(handle, tmp_path / "f.lsh")
    assert jx.hash_input() == (True, oracle.lsh_scale_exp(bf))
    for x, y in zip(ref, jx.query(qf)):
        np.testing.assert_array_equal(x, y)
    jx.close()
    (tmp_path / "bad.lsh").write_bytes(b"not an index")
    with pytest.raises(SiftError):
        LSHIndex.load(handle, tmp_path / "bad.lsh")
    data = (tmp_path / "ix12.lsh").read_bytes()
    (tmp_path / "cut.lsh").write_bytes(data[:-7])
    with pytest.raises(SiftError):
        LSHIndex.load(handle, tmp_path / "cut.lsh")


def smallest_bits(margins, P):
    """Bits of the P smallest margins, lowest bit first on ties ([..., bits])."""
    return np.argsort(margins, axis=-1, kind="stable")[..., :P]


# (12, 5, 12) and (24, 5, 13): more than 64 (table, probe) lookups per query,
# through the offset table and through the binary search
@pytest.mark.parametrize("nbits,L,P", [(12, 2, 3), (16, 1, 5), (24, 2, 2), (12, 5, 12), (24, 5, 13)])
def test_lsh_multiprobe_matches_oracle(handle, nbits, L, P):
    """Multi-probe: per table the base code, then the codes with each of the P
    bits of smallest |h.x| flipped (exact integers here, so ties go to the
    lower bit); buckets, candidate counts and re-ranked results equal the
    oracle's LSH query over those codes."""
    b = datagen.base(12_000, seed=33)
    q = datagen.queries(b, 300, seed=34)
    ix = LSHIndex(handle, b, nbits, L, seed=9)
    ix.set_probes(P)
    say("multiprobe", nbits, L, P)
    pc = ix.probe_codes(q)
    base = ix.codes(q)
    np.testing.assert_array_equal(pc[:, 0], base)
    H = ix.hyperplanes()
    for t in range(L):
        marg = np.abs(q.astype(np.float64) @ H[t].astype(np.float64).T)  # [Q][nbits], exact integers
        flips = pc[t, 1:] ^ base[t][None]  # [P][Q]: one bit each
        assert ((flips != 0) & ((flips & (flips - 1)) == 0)).all()
        np.testing.assert_array_equal(np.log2(flips).astype(np.int64).T, smallest_bits(marg, P))
    cb = ix.codes(b)
    cq = pc.reshape(L * (P + 1), -1)
    cr = np.repeat(cb, P + 1, axis=0)
    oi, od, _ = oracle.lsh_query(q, b, cq, cr)
    idx, dist, nc, mat = ix.query(q)
    np.testing.assert_array_equal(idx, oi)
    np.testing.assert_array_equal(dist, np.where(oi == 0xFFFFFFFF, FMAX, od.astype(np.float32)))
    want_nc = sum((cr[k][None, :] == cq[k][:, None]).sum(1) for k in range(cq.shape[0]))
    np.testing.assert_array_equal(nc, want_nc.astype(np.uint32))
    with pytest.raises(SiftError):
        ix.set_probes(min(nbits, 16) + 1)
    ix.close()


@pytest.mark.parametrize("nbits,L,P", [(8, 2, 0), (12, 5, 12), (16, 3, 2), (24, 5, 13), (32, 1, 0)])
def test_lsh_bucket_ordered_rerank(handle, monkeypatch, nbits, L, P):
    """The bucket-ordered re-rank (one wave per (table, probe, query) lookup in
    bucket order, then a per-query merge that drops rows found twice) gives
    the per-query re-rank's outputs exactly, and so the oracle's."""
    b = datagen.base(12_000, seed=41)
    q = datagen.queries(b, 500, seed=42)
    ix = LSHIndex(handle, b, nbits, L, seed=11)
    ix.set_probes(P)
    say("lsh order", nbits, L, P)
    monkeypatch.setenv("SIFT_LSH_ORDER", "0")
    want = ix.query(q)
    monkeypatch.setenv("SIFT_LSH_ORDER", "1")
    got = ix.query(q)
    for w, g in zip(want, got):
        np.testing.assert_array_equal(g, w)
    if P == 0:
        cq, cb = ix.codes(q), ix.codes(b)
        oi, _, _ = oracle.lsh_query(q, b, cq, cb)
        np.testing.assert_array_equal(got[0], oi)
    ix.close()


@pytest.mark.parametrize("nbits,L", [(4, 3), (12, 2), (24, 2), (32, 1)])
def test_lsh_bucket_stat(handle, nbits, L):
    """The index's mean bucket size seen by a base row equals the sum of
    squared bucket sizes over the tables / (L n), counted from the codes."""
    b = datagen.base(9_000, seed=51)
    ix = LSHIndex(handle, b, nbits, L, seed=13)
    say("lsh bucket stat", nbits, L)
    cb = ix.codes(b)
    want = sum(float((np.unique(cb[t], return_counts=True)[1].astype(np.float64) ** 2).sum())
               for t in range(L)) / (L * len(b))
    assert ix.bucket_stat() == pytest.approx(want, rel=1e-12)
    ix.close()


@pytest.mark.parametrize("nbits", [12, 24])
@pytest.mark.parametrize("distinct", [1, 40])
def test_lsh_bucket_stat_large_buckets(handle, nbits, distinct):
    """Buckets far larger than the statistic's 4096-position blocks (rows
    repeated from a few distinct ones, or one row only)."""
    src = datagen.base(distinct, seed=52)
    b = np.ascontiguousarray(src[np.random.default_rng(5).integers(0, distinct, 13_000)])
    ix = LSHIndex(handle, b, nbits, 2, seed=17)
    say("lsh bucket stat, large buckets", nbits, distinct)
    cb = ix.codes(b)
    want = sum(float((np.unique(cb[t], return_counts=True)[1].astype(np.float64) ** 2).sum())
               for t in range(2)) / (2 * len(b))
    assert ix.bucket_stat() == pytest.approx(want, rel=1e-12)
    ix.close()


@pytest.mark.parametrize("seed", [3, 17])
def test_lsh_random_configs(handle, monkeypatch, seed):
    """Random (nbits, L, probes, hyperplane kind) with base and query sizes at
    wave and tile edges, natural or 3-letter-alphabet data (many equal
    projections and distances): codes equal the oracle's, and both re-rank
    schedules return the oracle's LSH query over the index's probe codes."""
    rng = np.random.default_rng(seed)
    for it in range(16):
        nbits, L = int(rng.integers(1, 33)), int(rng.integers(1, 7))
        while nbits + int(np.ceil(np.log2(L))) > 32:  # table number shares the 32-bit key
            L -= 1
        P = int(rng.integers(0, min(nbits, 16) + 1)) if rng.random() < 0.5 else 0
        n = int(rng.choice([1, 2, 31, 32, 33, 700, 5000]))
        qn = int(rng.choice([1, 7, 64, 65, 300]))
        ties = bool(rng.integers(0, 2))
        hk = LSH_HYPERPLANE_PM1 if rng.random() < 0.3 else LSH_HYPERPLANE
        if ties:
            b = (rng.integers(0, 3, size=(n, 128)) * 85).astype(np.float32)
            q = (rng.integers(0, 3, size=(qn, 128)) * 85).astype(np.float32)
        else:
            b = datagen.base(n, seed=int(rng.integers(1 << 30)))
            q = datagen.queries(b, qn, seed=int(rng.integers(1 << 30)))
        say("lsh random", seed, it, "nbits", nbits, "L", L, "P", P, "n", n, "q", qn, "ties", ties, "kind", hk)
        ix = LSHIndex(handle, b, nbits, L, hash_kind=hk, seed=int(rng.integers(1 << 20)))
        ix.set_probes(P)
        cb = ix.codes(b)
        np.testing.assert_array_equal(cb, oracle_tables(b, ix.hyperplanes()))
        pc = ix.probe_codes(q) if P else ix.codes(q)[:, None]
        cq = pc.reshape(L * (P + 1), -1)
        oi, od, _ = oracle.lsh_query(q, b, cq, np.repeat(cb, P + 1, axis=0))
        for order in ("0", "1"):
            monkeypatch.setenv("SIFT_LSH_ORDER", order)
            idx, dist, _, _ = ix.query(q)
            np.testing.assert_array_equal(idx, oi, err_msg="order %s" % order)
            np.testing.assert_array_equal(dist, np.where(oi == 0xFFFFFFFF, FMAX, od.astype(np.float32)))
        ix.close()


def test_lsh_float_fallback(handle):
    """A float-mode index with an F32 brute-force index as its exact fallback:
    queries falling back get host_brute's answer (coarse keys, exact re-rank
    from the F32 tiles); the rest keep the float LSH answer."""
    from matching_sift_amd import COMPUTE_F32, BruteForceIndex
    b = datagen.unit(6000, seed=91)
    rng = np.random.default_rng(92)
    q = (b[rng.integers(0, 6000, 120)] + rng.normal(0, 2e-3, (120, 128))).astype(np.float32)
    q = np.concatenate([q, datagen.unit(80, seed=93)])
    ix = LSHIndex(handle, b, 16, 2, seed=4)
    bf = BruteForceIndex(handle, b, compute=COMPUTE_F32)
    say("lsh float fallback")
    lsh_idx, lsh_dist, nc, _ = ix.query(q)
    idx, dist, nc2, mat, nfb = ix.query_fallback(q, bf, min_cand=20)
    s = oracle.host_brute(q, b, threads=8)
    low = nc < 20
    assert nfb == int(low.sum()) and 0 < nfb < len(q)
    np.testing.assert_array_equal(idx[low], s[low, 2:4].astype(np.uint32))
    np.testing.assert_array_equal(dist[low], s[low, 0:2])
    np.testing.assert_array_equal(idx[~low], lsh_idx[~low])
    np.testing.assert_array_equal(dist[~low], lsh_dist[~low])
    bf.close()
    ix.close()
