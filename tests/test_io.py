"""SIFT vector files (.fvecs / .ivecs / .bvecs, as src/test.cu:76-155 reads
them) and the recall count of src/test.cu:224-301, through the C ABI.  The
file tests and the recall semantics run on the CPU; the last test checks the
GPU brute force against a ground-truth file written from an exact ranking."""
import os
import sys

import numpy as np
import pytest

from matching_sift_amd import SiftError, datagen, read_vecs, recall_count, vecs_info


def write_vecs(path, rows, dtype):
    rows = np.asarray(rows)
    with open(path, "wb") as f:
        for r in rows:
            f.write(np.int32(len(r)).tobytes())
            f.write(np.asarray(r, dtype).tobytes())


def test_fvecs_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    x = rng.integers(0, 256, size=(37, 128)).astype(np.float32)
    p = str(tmp_path / "base.fvecs")
    write_vecs(p, x, np.float32)
    assert vecs_info(p) == (37, 128)
    np.testing.assert_array_equal(read_vecs(p), x)
    np.testing.assert_array_equal(read_vecs(p, n=5, first=30), x[30:35])
    # the reference's real-data mode divides by 512
    np.testing.assert_array_equal(read_vecs(p, scale=1 / 512), x * np.float32(1 / 512))
    np.testing.assert_array_equal(read_vecs(p, cols=16), x[:, :16])


def test_bvecs_and_ivecs(tmp_path):
    rng = np.random.default_rng(2)
    b = rng.integers(0, 256, size=(9, 128)).astype(np.uint8)
    pb = str(tmp_path / "x.bvecs")
    write_vecs(pb, b, np.uint8)
    np.testing.assert_array_equal(read_vecs(pb), b.astype(np.float32))
    gt = np.stack([rng.permutation(1000)[:100] for _ in range(6)]).astype(np.int32)
    pi = str(tmp_path / "gt.ivecs")
    write_vecs(pi, gt, np.int32)
    g = read_vecs(pi, cols=10)  # first `recall` of 100 ids
    assert g.dtype == np.int32
    np.testing.assert_array_equal(g, gt[:, :10])
    np.testing.assert_array_equal(read_vecs(pi, n=2, first=4), gt[4:6])


def test_bad_files(tmp_path):
    p = str(tmp_path / "t.fvecs")
    write_vecs(p, np.zeros((3, 8)), np.float32)
    with open(p, "ab") as f:
        f.write(b"\x01\x02")  # not a whole record
    with pytest.raises(SiftError):
        vecs_info(p)
    q = str(tmp_path / "mixed.fvecs")
    with open(q, "wb") as f:  # same size as two 4-dim records, second header says 5
        f.write(np.int32(4).tobytes() + np.zeros(4, np.float32).tobytes())
        f.write(np.int32(5).tobytes() + np.zeros(4, np.float32).tobytes())
    assert vecs_info(q) == (2, 4)
    with pytest.raises(SiftError):
        read_vecs(q)
    with pytest.raises(SiftError):
        read_vecs(str(tmp_path / "missing.fvecs"))
    ok = str(tmp_path / "ok.fvecs")
    write_vecs(ok, np.zeros((2, 4)), np.float32)
    with pytest.raises(SiftError):
        read_vecs(ok, n=3)
    with pytest.raises(SiftError):
        read_vecs(ok, cols=5)


def test_recall_count_semantics():
    """Constructed answers: query q's match sits at position pos[q] of its
    ground-truth list (or is absent / UINT32_MAX); found at depth R exactly
    when pos[q] < R."""
    rng = np.random.default_rng(3)
    nq, k = 200, 100
    gt = np.stack([rng.permutation(10_000)[:k] for _ in range(nq)]).astype(np.int32)
    pos = rng.integers(0, 130, size=nq)  # >= 100: not in the list
    found = np.array([gt[q, p] if p < k else 10_000 + q for q, p in enumerate(pos)], np.uint32)
    found[::17] = 0xFFFFFFFF
    pos[::17] = k
    depths = (1, 10, 20, 50, 100)
    got = recall_count(found, gt, depths)
    for R in depths:
        assert got[R] == int((pos < R).sum())
    with pytest.raises(SiftError):
        recall_count(found, gt, (0,))
    with pytest.raises(SiftError):
        recall_count(found, gt, (101,))


def exact_topk(q, b, k):
    """Exact ranking by integer squared L2, lowest index first on ties."""
    qi, bi = q.astype(np.int64), b.astype(np.int64)
    d = (qi * qi).sum(1)[:, None] + (bi * bi).sum(1)[None, :] - 2 * qi @ bi.T
    return np.argsort(d, axis=1, kind="stable")[:, :k].astype(np.int32)


def test_ground_truth_file_vs_oracle(tmp_path):
    """A ground-truth file written from an exact ranking agrees with the
    oracle's 2-NN: recall@1 of its NN1 and recall@2 of its NN2 are complete."""
    import oracle
    b = datagen.base(3000, seed=11)
    q = datagen.queries(b, 60, seed=12)
    p = str(tmp_path / "gt.ivecs")
    write_vecs(p, exact_topk(q, b, 100), np.int32)
    gt = read_vecs(p)
    s = oracle.host_brute(q, b, threads=4)
    assert recall_count(s[:, 2].astype(np.uint32), gt, (1,))[1] == len(q)
    assert recall_count(s[:, 3].astype(np.uint32), gt[:, 1:], (1,))[1] == len(q)


@pytest.mark.gpu
def test_gpu_recall_against_ground_truth_file(handle, tmp_path):
    """The GPU brute force against a ground-truth file: every first neighbour
    is the file's first id (recall@1 complete), as the reference's harness
    counts it."""
    from matching_sift_amd import COMPUTE_BF16, knn2
    b = datagen.base(20_000, seed=13)
    q = datagen.queries(b, 200, seed=14)
    fb, fq, fg = (str(tmp_path / n) for n in ("sift_base.fvecs", "sift_query.fvecs", "sift_groundtruth.ivecs"))
    write_vecs(fb, b, np.float32)
    write_vecs(fq, q, np.float32)
    write_vecs(fg, exact_topk(q, b, 100), np.int32)
    print("gpu recall vs ground-truth file", flush=True)
    sys.stdout.flush()
    idx = knn2(handle, read_vecs(fq), read_vecs(fb), compute=COMPUTE_BF16)[0]
    hits = recall_count(idx[:, 0], read_vecs(fg))
    assert hits[1] == len(q) and hits[100] == len(q)
