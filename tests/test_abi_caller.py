"""A compiled C++ caller in the shape of the reference's driver
(tests/abi_caller.cpp, after src/test.cu:45-215) builds against
include/sift2nn.h and links libsift2nn.so (CPU tier), and its outputs equal
the oracle's on the GPU: des_t_f buffers, a float4 output, device_brute,
cublas_2nn_f, cublas_2nn_sift with types 0 and 1, and lsh_gpu."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "abi_caller.cpp")
BIN = os.path.join(ROOT, "tests", "_build", "abi_caller")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build_caller():
    """Compile the caller against the public header only and link the
    product library (found next to the binary's tree at run time)."""
    from matching_sift_amd import LIB_PATH
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), SRC,
           "-o", BIN + ".tmp", "-L", os.path.dirname(LIB_PATH), "-lsift2nn",
           "-Wl,-rpath,$ORIGIN/../../matching_sift_amd"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    os.replace(BIN + ".tmp", BIN)
    return BIN


def test_caller_compiles_and_links():
    """The reference caller's types compile against the header: des_t_f* into
    const des_t_f*, float4* into the sorted output, sift_handle_t for the
    cuBLAS handle; and every call resolves in libsift2nn.so."""
    b = build_caller()
    out = subprocess.run(["nm", "-u", b], stdout=subprocess.PIPE, text=True, check=True).stdout
    for name in ("device_brute", "cublas_2nn_f", "cublas_2nn_sift", "lsh_gpu", "sift_create"):
        assert name in out
    assert "_Z" not in "".join(l for l in out.splitlines() if "sift" in l or "brute" in l or "lsh" in l)


@pytest.mark.gpu
def test_caller_outputs_equal_the_oracle(handle, tmp_path):
    from matching_sift_amd import LSHIndex, datagen
    if not os.path.exists(BIN) or os.path.getmtime(BIN) < os.path.getmtime(SRC):
        build_caller()
    b = datagen.base(12_000, seed=601)
    q = datagen.queries(b, 700, seed=602)
    q.tofile(tmp_path / "q.f32")
    b.tofile(tmp_path / "r.f32")
    seed = 4242
    print("running abi_caller", flush=True)
    r = subprocess.run([BIN, str(tmp_path / "q.f32"), str(tmp_path / "r.f32"), "700", "12000", str(seed),
                        str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout
    s = oracle.host_brute(q, b, threads=8)
    for name in ("device_brute.f4", "cublas_2nn_f.f4"):
        got = np.fromfile(tmp_path / name, np.float32).reshape(700, 4)
        np.testing.assert_array_equal(got, s)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "sift_type0.u32", np.uint32), oracle.ratio(s, 2.0))
    np.testing.assert_array_equal(np.fromfile(tmp_path / "sift_type1.u32", np.uint32), oracle.ratio(s, 0.8))
    # lsh_gpu: the same hyperplanes as an index built from the caller's seed
    ix = LSHIndex(handle, b, 12, 4, seed=seed)
    H = ix.hyperplanes()
    ix.close()
    cr = np.stack([oracle.lsh_codes(b, H[t]) for t in range(4)])
    cq = np.stack([oracle.lsh_codes(q, H[t]) for t in range(4)])
    oi, od, _ = oracle.lsh_query(q, b, cq, cr)
    fmax = np.float32(np.finfo(np.float32).max)
    d = np.where(oi == 0xFFFFFFFF, fmax, od.astype(np.float32))
    s4 = np.concatenate([d, np.where(oi == 0xFFFFFFFF, fmax, oi.astype(np.float32))], 1)
    want = oracle.ratio(s4, 0.8)
    want[oi[:, 0] == 0xFFFFFFFF] = 0xFFFFFFFF
    np.testing.assert_array_equal(np.fromfile(tmp_path / "lsh.u32", np.uint32), want)
