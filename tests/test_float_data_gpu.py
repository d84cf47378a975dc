"""GPU parity on descriptors that are not integers in [0,255]: the
reference's own test data (unit-norm floats, src/helper.cu:40-58) and real
SIFT divided by 512 (src/test.cu:103), with near-duplicate queries (d << 1).

The float kinds rank with the MFMAs (values rounded to the kind), then the 16
best coarse candidates of each query are re-ranked with host_lenght's
sequential FP32 sum (src/knn_brute.cu:873-885).  Checked against the oracle
(host_brute): idx1, idx2, d1, d2 and the ratio test.

Tolerance: FP32 and FP16 must equal host_brute exactly on these inputs.  BF16
rounds each value to 8 significant bits, so its coarse pass can rank a row
whose distance is within 2^-6 (|q| + |r|)^2 of the true neighbour's above it;
such a swap is allowed (and only such a swap): the returned row's exact
distance must lie within that bound of the oracle's, and every returned
distance must be host_lenght of the returned row.
"""
import sys

import numpy as np
import pytest

import oracle
from matching_sift_amd import (BF16_DEVICE, COMPUTE_AUTO, COMPUTE_BF16, COMPUTE_F16, COMPUTE_F32, COMPUTE_I8,
                               FLOAT_DEVICE, HALF_DEVICE, BruteForceIndex, DeviceArray, SiftError, api, datagen,
                               knn2)

pytestmark = pytest.mark.gpu
FMAX = np.float32(np.finfo(np.float32).max)
REL = {COMPUTE_BF16: 2.0 ** -6, COMPUTE_F16: 2.0 ** -9, COMPUTE_F32: 2.0 ** -16, COMPUTE_AUTO: 2.0 ** -16}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def near_duplicates(r, m, sd, seed, fresh=None):
    """m noisy copies of random rows of r (Normal(0, sd) per value) and
    `fresh` unrelated rows (unit-norm)."""
    rng = np.random.default_rng(seed)
    q = (r[rng.integers(0, r.shape[0], m)] + rng.normal(0, sd, (m, 128))).astype(np.float32)
    if fresh:
        scale = np.float32(np.sqrt((r.astype(np.float64) ** 2).sum(1)).mean())
        q = np.concatenate([q, datagen.unit(fresh, seed=seed + 1) * scale])
    return q


def check(q, r, srt, compute, want=None):
    """srt (float4 rows) against host_brute within the stated tolerance."""
    s = oracle.host_brute(q, r, threads=8) if want is None else want
    if compute != COMPUTE_BF16:
        np.testing.assert_array_equal(srt, s)
        return s
    diff = np.nonzero((srt != s).any(1))[0]
    say("bf16: %d of %d queries differ from host_brute" % (len(diff), q.shape[0]))
    assert len(diff) <= max(2, q.shape[0] // 50)
    nq = np.sqrt((q.astype(np.float64) ** 2).sum(1))
    nr = np.sqrt((r.astype(np.float64) ** 2).sum(1))
    for i in diff:
        for k in (0, 1):
            j = int(srt[i, 2 + k])
            assert j < r.shape[0]
            assert srt[i, k] == oracle.host_lenght(q, r, [i], [j])[0]  # exact distance of the returned row
            if j != int(s[i, 2 + k]):
                jo = int(s[i, 2 + k])
                bound = REL[compute] * (nq[i] + max(nr[j], nr[jo])) ** 2
                assert abs(float(srt[i, k]) - float(s[i, k])) <= bound, (i, k, srt[i], s[i], bound)
    return s


@pytest.mark.parametrize("compute", [COMPUTE_F32, COMPUTE_F16, COMPUTE_BF16, COMPUTE_AUTO])
@pytest.mark.parametrize("n", [3000, 40_000])
def test_unit_norm_near_duplicates(handle, compute, n):
    """n = 40,000 takes the seeded launch chain (bounds widened by the kind's
    rounding bound, none lowered inside a launch)."""
    r = datagen.unit(n, seed=401)
    q = near_duplicates(r, 200, 1e-3, seed=402, fresh=100)
    say("unit-norm near duplicates", n, compute)
    idx, dist, srt, mat = knn2(handle, q, r, compute=compute, ratio=0.8)
    s = check(q, r, srt, compute)
    assert (s[:200, 0] < 0.01).all()  # the copies really are near duplicates
    if compute != COMPUTE_BF16:
        np.testing.assert_array_equal(mat, oracle.ratio(s, 0.8))
        np.testing.assert_array_equal(idx, srt[:, 2:4].astype(np.uint32))
        np.testing.assert_array_equal(dist, srt[:, 0:2])


@pytest.mark.parametrize("compute", [COMPUTE_F32, COMPUTE_F16, COMPUTE_BF16])
def test_sift_divided_by_512(handle, compute):
    """src/test.cu:103 feeds real SIFT / 512: SIFT-like integers scaled by
    2^-9 (exact in every kind, but not integers)."""
    b = datagen.base(40_000, seed=411) / np.float32(512)
    q = datagen.queries(b * 512, 300, seed=412) / np.float32(512)
    say("sift / 512", compute)
    srt = knn2(handle, q, b, compute=compute, ratio=0.8)[2]
    check(q, b, srt, compute)


@pytest.mark.parametrize("type_", [HALF_DEVICE, BF16_DEVICE])
def test_half_and_bf16_inputs(handle, type_):
    """Non-integer fp16 / bf16 inputs: the exact distances are over the stored
    values (converted to float exactly)."""
    r = datagen.unit(5000, seed=421)
    q = near_duplicates(r, 150, 5e-3, seed=422, fresh=50)
    if type_ == HALF_DEVICE:
        rr, qq = r.astype(np.float16), q.astype(np.float16)
        rf, qf = rr.astype(np.float32), qq.astype(np.float32)
    else:
        rr = (r.view(np.uint32) >> 16).astype(np.uint16)
        qq = (q.view(np.uint32) >> 16).astype(np.uint16)
        rf = (rr.astype(np.uint32) << 16).view(np.float32)
        qf = (qq.astype(np.uint32) << 16).view(np.float32)
    say("half/bf16 input", type_)
    for compute in (COMPUTE_F16, COMPUTE_F32):
        srt = knn2(handle, qq, rr, compute=compute, type_=type_)[2]
        check(qf, rf, srt, compute)


def test_reference_fp32_entry_points(handle):
    """device_brute and cublas_2nn_f (src/knn_brute.cu:619, :523) on the
    reference's unit-norm data: host_brute exactly."""
    r = datagen.unit(20_000, seed=431)
    q = near_duplicates(r, 300, 1e-3, seed=432, fresh=100)
    s = oracle.host_brute(q, r, threads=8)
    qd, rd = DeviceArray.from_numpy(q), DeviceArray.from_numpy(r)
    out = DeviceArray.empty((q.shape[0], 4), np.float32)
    say("device_brute unit-norm")
    api.device_brute(qd, rd, q.shape[0], r.shape[0], out)
    np.testing.assert_array_equal(out.numpy(), s)
    say("cublas_2nn_f unit-norm")
    out.fill(0)
    api.cublas_2nn_f(qd, rd, q.shape[0], r.shape[0], out, handle)
    np.testing.assert_array_equal(out.numpy(), s)


def test_brute_force_index_float_data(handle):
    """An F32 index keeps the base exactly (its packed tiles) and re-ranks from
    them; BF16 and F16 indexes over integer bases re-rank float queries from
    their tiles too; a BF16 index over a float base is refused."""
    r = datagen.unit(30_000, seed=441)
    q = near_duplicates(r, 200, 1e-3, seed=442, fresh=60)
    say("F32 index, unit-norm base")
    ix = BruteForceIndex(handle, r, compute=COMPUTE_F32)
    s = oracle.host_brute(q, r, threads=8)
    for part in (slice(0, 260), slice(7, 40)):
        idx, dist = ix.query(q[part])
        np.testing.assert_array_equal(idx, s[part, 2:4].astype(np.uint32))
        np.testing.assert_array_equal(dist, s[part, 0:2])
    ix.close()
    with pytest.raises(SiftError) as e:
        BruteForceIndex(handle, r, compute=COMPUTE_BF16)
    assert e.value.code == -5
    b = datagen.base(30_000, seed=443)
    qb = datagen.queries(b, 200, seed=444) + np.float32(0.25)  # integer base, non-integer queries
    sb = oracle.host_brute(qb, b, threads=8)
    for compute in (COMPUTE_F16, COMPUTE_BF16, COMPUTE_F32):
        say("index over integer base, float queries", compute)
        ix = BruteForceIndex(handle, b, compute=compute)
        idx, dist = ix.query(qb)
        got = np.concatenate([dist, idx.astype(np.float32)], 1)
        check(qb, b, got, compute, want=sb)
        ix.close()
    ix = BruteForceIndex(handle, b, compute=COMPUTE_I8)
    with pytest.raises(SiftError) as e:
        ix.query(qb)
    assert e.value.code == -5
    ix.close()


def test_integer_data_after_float_data(handle, config0):
    """A call on integer data after float calls on the same handle is exact
    again (the coarse flag is per call)."""
    r = datagen.unit(2000, seed=451)
    knn2(handle, r[:10] + np.float32(1e-3), r, compute=COMPUTE_BF16)
    q, b = config0
    srt = knn2(handle, q, b, compute=COMPUTE_BF16)[2]
    np.testing.assert_array_equal(srt, oracle.host_brute(q, b, threads=8))
