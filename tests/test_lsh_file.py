"""sift_lsh_check_file (and sift_lsh_load, which runs the same checks) on the
host, no device: a well-formed index file passes, and every kind of damage
that would make the GPU gather outside the base or look up the wrong bucket
is refused before anything reaches the device."""
import struct

import numpy as np
import pytest

from matching_sift_amd import SiftError, lsh_check_file

N, L, NBITS = 6, 2, 3


def write_index(path, keys=None, vals=None, u8=None, nrm=None, planes=None, header=None, trunc=0):
    rng = np.random.default_rng(7)
    if planes is None:
        planes = rng.integers(-127, 128, size=(L, NBITS, 128)).astype(np.float32)
    if u8 is None:
        u8 = rng.integers(0, 256, size=(N, 128)).astype(np.uint8)
    if nrm is None:
        nrm = (u8.astype(np.int64) ** 2).sum(1).astype(np.int32)
    if keys is None:
        codes = rng.integers(0, 1 << NBITS, size=(L, N)).astype(np.uint32)
        keys, vals = [], []
        for t in range(L):
            order = np.lexsort((np.arange(N), codes[t]))
            keys.append((t << NBITS) | codes[t][order])
            vals.append(order.astype(np.uint32))
        keys, vals = np.concatenate(keys).astype(np.uint32), np.concatenate(vals).astype(np.uint32)
    hd = header or dict(version=1, n=N, nbits=NBITS, L=L, kind=0, nparam=0, seed=5)
    blob = struct.pack("<8sIIiiiIQ", b"SIFTLSH1", hd["version"], hd["n"], hd["nbits"], hd["L"], hd["kind"],
                       hd["nparam"], hd["seed"])
    blob += planes.tobytes() + keys.tobytes() + vals.tobytes() + u8.tobytes() + nrm.tobytes()
    with open(path, "wb") as f:
        f.write(blob[:len(blob) - trunc])
    return keys, vals, u8, nrm


def refused(path):
    with pytest.raises(SiftError) as e:
        lsh_check_file(path)
    assert e.value.code == -1
    return str(e.value)


def test_well_formed_file_passes(tmp_path):
    p = tmp_path / "ok.lsh"
    write_index(p)
    assert lsh_check_file(p) == 0


def test_damaged_files_are_refused(tmp_path):
    p = tmp_path / "x.lsh"
    keys, vals, u8, nrm = write_index(p)
    v = vals.copy()
    v[3] = N  # a row past the base: the re-rank would read outside it
    write_index(p, keys=keys, vals=v)
    assert "past the base" in refused(p)
    k = keys.copy()
    k[0], k[1] = (1 << NBITS) - 1, 0  # table 0 out of order
    write_index(p, keys=k, vals=vals)
    assert "not sorted" in refused(p)
    k = keys.copy()
    k[N] = k[N] & ((1 << NBITS) - 1)  # table 1's first key without its table prefix
    write_index(p, keys=k, vals=vals)
    assert "outside its table" in refused(p)
    n2 = nrm.copy()
    n2[0] += 1
    write_index(p, keys=keys, vals=vals, u8=u8, nrm=n2)
    assert "norm" in refused(p)
    planes = np.zeros((L, NBITS, 128), np.float32)
    planes[0, 0, 0] = 0.5
    write_index(p, keys=keys, vals=vals, planes=planes)
    assert "hyperplane" in refused(p)
    write_index(p, trunc=4)
    assert "size" in refused(p)
    write_index(p, header=dict(version=1, n=N, nbits=40, L=L, kind=0, nparam=0, seed=5))
    assert "header" in refused(p)
    with open(p, "wb") as f:
        f.write(b"not an index")
    refused(p)
    with pytest.raises(SiftError):
        lsh_check_file(tmp_path / "missing.lsh")
