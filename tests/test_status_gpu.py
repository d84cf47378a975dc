"""Failure reporting on the GPU, forced from the host so every kernel still
exits normally.

- The 8-wave tile shapes hand LDS stages between waves with bounded waits.
  With the poll limit set to 0 (SIFT_WAIT_POLLS at handle creation) a wave
  that finds a stage not yet ready gives up at once; the launch still ends,
  and the blocking call must report SIFT_ERR_HIP instead of returning
  results it cannot vouch for.  A handle with the default limit is exact.
- sift_lsh_query_fallback must report non-integer queries like
  sift_lsh_query does.
"""
import sys

import numpy as np
import pytest

import oracle
from matching_sift_amd import (COMPUTE_BF16, COMPUTE_I8, BruteForceIndex, DeviceArray, Handle, LSHIndex,
                               PairsPlan, SiftError, datagen, knn2)

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_wait_give_up_is_reported(monkeypatch):
    b = datagen.base(20_000, seed=501)
    q = datagen.queries(b, 600, seed=502)
    monkeypatch.setenv("SIFT_SHAPE", "2,8,2")
    monkeypatch.setenv("SIFT_WAIT_POLLS", "0")
    h = Handle(seed=1)
    try:
        say("knn2 with 0 polls")
        with pytest.raises(SiftError) as e:
            knn2(h, q, b, compute=COMPUTE_BF16)
        assert e.value.code == -3 and "gave up" in str(e.value)
        say("pairs with 0 polls")
        q_off = np.array([0, 600], np.uint32)
        r_off = np.array([0, 20_000], np.uint32)
        plan = PairsPlan(h, q_off, r_off, np.array([[0, 0]], np.uint32), compute=COMPUTE_BF16, flags=0)
        plan.set_shape(2, 8, 2)
        with pytest.raises(SiftError) as e:
            plan.execute(DeviceArray.from_numpy(q), DeviceArray.from_numpy(b))
        assert e.value.code == -3
        plan.close()
    finally:
        h.close()
    monkeypatch.delenv("SIFT_WAIT_POLLS")
    h = Handle(seed=1)
    try:
        say("knn2 with the default poll limit")
        srt = knn2(h, q, b, compute=COMPUTE_BF16)[2]
        np.testing.assert_array_equal(srt, oracle.host_brute(q, b, threads=8))
    finally:
        h.close()


def test_lsh_fallback_reports_non_integer_queries(handle):
    b = datagen.base(5000, seed=511)
    ix = LSHIndex(handle, b, 12, 2, seed=3)
    bf = BruteForceIndex(handle, b, compute=COMPUTE_I8)
    q = datagen.queries(b, 50, seed=512) + np.float32(0.5)
    say("lsh fallback, non-integer queries")
    with pytest.raises(SiftError) as e:
        ix.query_fallback(q, bf)
    assert e.value.code == -5
    qi = datagen.queries(b, 50, seed=513)
    idx, dist = ix.query_fallback(qi, bf, min_cand=10**9)[:2]  # every query falls back: exact
    s = oracle.host_brute(qi, b, threads=8)
    np.testing.assert_array_equal(idx, s[:, 2:4].astype(np.uint32))
    bf.close()
    ix.close()
