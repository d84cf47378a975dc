"""GPU parity of the learned hash (no counterpart in the reference): the
HIP forward against the oracle's quantised forward, the HIP backward
(loss, straight-through gradient, weight gradients) against the oracle's
float64 backward, the trained network against the untrained one, and the
learned LSH index and exact fallback against the oracle's LSH query."""
import sys

import numpy as np
import pytest

import oracle
from matching_sift_amd import (BruteForceIndex, DeviceArray, HashMLP, LSHIndex, SiftError, datagen, knn2,
                               COMPUTE_BF16)

pytestmark = pytest.mark.gpu
W = dict(alpha=0.7, beta=0.9, gamma=0.3)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def random_params(rng, nbits):
    p = np.zeros(oracle.HASH_NPARAM, np.float32)
    p[:256 * 128] = rng.normal(0, 1.28, 256 * 128)  # stored weights; layer 1 uses them / 64
    p[256 * 128:256 * 129] = rng.normal(0, 0.5, 256)
    w2 = rng.normal(0, 0.06, (32, 256)).astype(np.float32)
    w2[nbits:] = 0
    p[256 * 129:256 * 129 + 32 * 256] = w2.ravel()
    p[256 * 129 + 32 * 256:256 * 129 + 32 * 256 + nbits] = rng.normal(0, 0.1, nbits)
    return p


def codes_of(z, nbits):
    bits = (z[:, :nbits] < 0).astype(np.uint64)
    return (bits << np.arange(nbits, dtype=np.uint64)).sum(1).astype(np.uint32)


@pytest.mark.parametrize("nbits", [16, 24, 32])
def test_forward_matches_oracle(handle, nbits):
    rng = np.random.default_rng(nbits)
    m = HashMLP(handle, nbits, seed=5)
    p = random_params(rng, nbits)
    m.set_params(p)
    np.testing.assert_array_equal(m.params(), p)
    x = datagen.base(300, seed=40 + nbits)
    say("forward", nbits)
    codes, z = m.codes(x, with_z=True)
    _, zo = oracle.hash_forward(p, nbits, x, quant=True)
    scale = np.abs(zo).max()
    np.testing.assert_allclose(z, zo, rtol=1e-3, atol=1e-4 * scale)
    safe = (np.abs(zo[:, :nbits]) > 1e-4 * scale).all(1)  # no output within rounding of 0
    assert safe.mean() > 0.9
    np.testing.assert_array_equal(codes[safe], codes_of(zo, nbits)[safe])
    m.close()


def test_gradient_matches_oracle(handle):
    """One batch of 16 pairs: the loss and every parameter gradient of the HIP
    backward against the oracle (float64, same roundings), relative to the
    gradient's scale.  The batch is the first seed whose outputs all keep a
    margin of 1e-3 from the sign and straight-through thresholds (0 and +-1),
    so FP32 and float64 agree on every sign and mask."""
    nbits, B = 24, 16
    for seed in range(3, 200):
        rng = np.random.default_rng(seed)
        p = random_params(rng, nbits)
        b = datagen.base(4 * B, seed=seed + 1000)
        x, y = b[:B], datagen.queries(b[:B], B, seed=seed + 2000, frac_copy=1.0)
        xy = np.concatenate([x, y])
        hid, zo = oracle.hash_forward(p, nbits, xy, quant=True)
        za = np.abs(zo[:, :nbits])
        if za.min() > 1e-3 and np.abs(za - 1).min() > 1e-3:
            break
    else:
        pytest.fail("no batch with margins")
    m = HashMLP(handle, nbits)
    m.set_params(p)
    say("gradient batch, seed", seed)
    loss, g, z = m.grad(x, y, **W)
    scale = np.abs(zo).max()
    # a hidden value whose FP32 and float64 sums round to neighbouring
    # bfloat16 values moves z by one hidden ulp times a W2 weight
    np.testing.assert_allclose(z, zo, rtol=1e-3, atol=1e-4 * scale)
    L, dls, dz = oracle.hash_loss_grad(zo, B, nbits, W["alpha"], W["beta"], W["gamma"])
    assert abs(loss[0] - L) <= 1e-4 * max(1.0, abs(L))
    go = oracle.hash_backward(p, nbits, xy, hid, dz, quant=True)
    for name, sl in (("W1", slice(0, 256 * 128)), ("b1", slice(256 * 128, 256 * 129)),
                     ("W2", slice(256 * 129, 256 * 129 + 32 * 256)), ("b2", slice(256 * 129 + 32 * 256, None))):
        gs, os_ = g[sl].astype(np.float64), go[sl]
        tol = 2e-3 * np.abs(os_).max() + 1e-12
        err = np.abs(gs - os_).max()
        say("grad", name, "max |err|", err, "scale", np.abs(os_).max())
        assert err <= tol, name
    m.close()


def test_training_gains_over_untrained(handle):
    """Trained on exact 2-NN labels, the network is a more selective hash than
    the same network untrained and than random hyperplanes of the same bits:
    selectivity = P(query and its true nearest row share a bucket) divided by
    the fraction of the base in the query's bucket (a random row's chance).
    Its loss falls."""
    nbits = 16
    b = datagen.base(30_000, seed=71)
    xtr = datagen.queries(b, 8_000, seed=72)
    xte = datagen.queries(b, 2_000, seed=73)
    nn = knn2(handle, xte, b, compute=COMPUTE_BF16)[0][:, 0]
    untrained = HashMLP(handle, nbits, seed=9)
    trained = HashMLP(handle, nbits, seed=9)
    say("training")
    el = trained.train(xtr, b, epochs=8, batch=1024, lr=1e-3, seed=1, **W)
    say("epoch losses", el)
    assert el[-1] < el[0]

    def stats(cq, cb):
        # fraction of queries sharing their true nearest row's bucket, and the
        # mean bucket size a query lands in (the candidates it would examine)
        u, cnt = np.unique(cb, return_counts=True)
        load = dict(zip(u.tolist(), cnt.tolist()))
        return float((cq == cb[nn]).mean()), float(np.mean([load.get(int(c), 0) for c in cq]))
    hu, lu = stats(untrained.codes(xte), untrained.codes(b))
    ht, lt = stats(trained.codes(xte), trained.codes(b))
    ix = LSHIndex(handle, b, nbits, 1, seed=5)
    hp = ix.codes(xte)[0], ix.codes(b)[0]
    hh, lh = stats(*hp)
    say("same-bucket rate / bucket load: untrained %.3f / %.1f, trained %.3f / %.1f, hyperplanes %.3f / %.1f"
        % (hu, lu, ht, lt, hh, lh))
    sel = lambda hit, load: hit * len(b) / max(load, 1e-9)  # noqa: E731
    assert ht > 0.02
    assert sel(ht, lt) > 2 * sel(hu, lu) and sel(ht, lt) > sel(hh, lh)
    ix.close()
    untrained.close()
    trained.close()


def test_learned_lsh_matches_oracle(handle):
    """Buckets and re-ranked results of the learned index equal the oracle's
    LSH query given the networks' codes."""
    nbits, L = 12, 2
    b = datagen.base(5_000, seed=81)
    q = datagen.queries(b, 700, seed=82)
    mlps = [HashMLP(handle, nbits, seed=100 + t) for t in range(L)]
    ix = LSHIndex.learned(handle, b, mlps)
    say("learned index query")
    idx, dist, nc, mat = ix.query(q)
    cq = np.stack([m.codes(q) for m in mlps])
    cb = np.stack([m.codes(b) for m in mlps])
    oi, od, _ = oracle.lsh_query(q, b, cq, cb)
    # candidates examined: bucket sizes summed over the tables
    want_nc = sum((cb[t][None, :] == cq[t][:, None]).sum(1) for t in range(L))
    np.testing.assert_array_equal(nc, want_nc.astype(np.uint32))
    np.testing.assert_array_equal(idx, oi)
    FMAX = np.float32(np.finfo(np.float32).max)
    np.testing.assert_array_equal(dist, np.where(od == 0xFFFFFFFF, FMAX, od.astype(np.float32)))
    ix.close()
    for m in mlps:
        m.close()


def test_fallback_is_exact(handle):
    """Queries with fewer than min_cand candidates are answered by the exact
    brute-force path; the others keep the LSH answer."""
    nbits, L = 16, 1
    b = datagen.base(8_000, seed=91)
    q = datagen.queries(b, 1_000, seed=92)
    mlp = HashMLP(handle, nbits, seed=7)
    ix = LSHIndex.learned(handle, b, [mlp])
    bf = BruteForceIndex(handle, b, compute=COMPUTE_BF16)
    li, ld, lnc, lm = ix.query(q)
    say("fallback query")
    idx, dist, nc, mat, nf = ix.query_fallback(q, bf, min_cand=32)
    np.testing.assert_array_equal(nc, lnc)
    low = lnc < 32
    assert nf == int(low.sum()) and 0 < nf < len(q)
    ei, ed, _, em = knn2(handle, q[low], b, compute=COMPUTE_BF16)
    np.testing.assert_array_equal(idx[low], ei)
    np.testing.assert_array_equal(dist[low], ed)
    np.testing.assert_array_equal(mat[low], em)
    np.testing.assert_array_equal(idx[~low], li[~low])
    # ratio rule: also the queries whose best candidate fails
    # sqrt(d1)/sqrt(d2) < 0.8 against the next one go to the exact path
    say("fallback query, ratio rule")
    idx, dist, nc, mat, nf = ix.query_fallback(q, bf, min_cand=32, accept_ratio=0.8)
    with np.errstate(divide="ignore", invalid="ignore"):
        amb = ~(np.sqrt(ld[:, 0]) / np.sqrt(ld[:, 1]) < np.float32(0.8))
    sel = low | amb
    assert nf == int(sel.sum()) and int(low.sum()) < nf < len(q)
    ei, ed, _, em = knn2(handle, q[sel], b, compute=COMPUTE_BF16)
    np.testing.assert_array_equal(idx[sel], ei)
    np.testing.assert_array_equal(dist[sel], ed)
    np.testing.assert_array_equal(mat[sel], em)
    np.testing.assert_array_equal(idx[~sel], li[~sel])
    np.testing.assert_array_equal(mat[~sel], lm[~sel])
    with pytest.raises(SiftError):
        other = BruteForceIndex(handle, b[:100], compute=COMPUTE_BF16)
        ix.query_fallback(q, other)
    with pytest.raises(SiftError):
        ix.query_fallback(q, bf, accept_ratio=1.5)
    ix.close()
    bf.close()
    mlp.close()


def test_learned_index_save_load(handle, tmp_path):
    """A learned index saved and loaded after its networks are gone (the
    loaded index owns copies) answers exactly as before."""
    b = datagen.base(6_000, seed=83)
    q = datagen.queries(b, 400, seed=84)
    mlps = [HashMLP(handle, 14, seed=300 + t) for t in range(2)]
    ix = LSHIndex.learned(handle, b, mlps)
    say("learned index save/load")
    ref = ix.query(q)
    ix.save(tmp_path / "learned.lsh")
    ix.close()
    for m in mlps:
        m.close()
    jx = LSHIndex.load(handle, tmp_path / "learned.lsh")
    assert (jx.n, jx.nbits, jx.L) == (len(b), 14, 2)
    for x, y in zip(ref, jx.query(q)):
        np.testing.assert_array_equal(x, y)
    jx.close()


def test_learned_multiprobe(handle):
    """Learned multi-probe: the probes flip the bits of smallest |z| (the
    network's own outputs, lower bit on ties) and the index answers as the
    oracle's LSH query over the probe codes; more probes find a query's
    nearest row at least as often."""
    nbits, L, P = 14, 2, 4
    b = datagen.base(6_000, seed=85)
    q = datagen.queries(b, 500, seed=86)
    mlps = [HashMLP(handle, nbits, seed=400 + t) for t in range(L)]
    ix = LSHIndex.learned(handle, b, mlps)
    base_idx = ix.query(q)[0]
    ix.set_probes(P)
    say("learned multiprobe")
    pc = ix.probe_codes(q)
    for t in range(L):
        codes, z = mlps[t].codes(q, with_z=True)
        np.testing.assert_array_equal(pc[t, 0], codes)
        flips = pc[t, 1:] ^ codes[None]
        assert ((flips != 0) & ((flips & (flips - 1)) == 0)).all()
        want = np.argsort(np.abs(z[:, :nbits]), axis=1, kind="stable")[:, :P]
        np.testing.assert_array_equal(np.log2(flips).astype(np.int64).T, want)
    cb = np.stack([m.codes(b) for m in mlps])
    cq = pc.reshape(L * (P + 1), -1)
    cr = np.repeat(cb, P + 1, axis=0)
    oi, od, _ = oracle.lsh_query(q, b, cq, cr)
    idx, dist, nc, mat = ix.query(q)
    np.testing.assert_array_equal(idx, oi)
    nn = knn2(handle, q, b, compute=COMPUTE_BF16)[0][:, 0]
    assert (idx[:, 0] == nn).mean() >= (base_idx[:, 0] == nn).mean()
    ix.close()
    for m in mlps:
        m.close()
