"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU tier):
the library's host sources (the SIFT vector reader and recall count, the LSH
index file checker, with every other host path linked in), the CPU oracle and
the synthetic data generator, driven by tests/sanitize_driver.cpp over valid,
damaged and byte-flipped files and small inputs.  The sanitizers instrument
host code only (-Xarch_host); no device code runs."""
import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "matching_sift_amd", "csrc")
OUT = os.path.join(ROOT, "tests", "_build", "sanitize")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANG = "/opt/rocm/llvm/bin/clang"
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]
HOST_SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]


def run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout[-4000:]


def build():
    os.makedirs(OUT, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = []
    for s in srcs:
        jobs.append([HIPCC, "-O1", "-g", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-Xclang", "-target-feature",
                     "-Xclang", "-packed-fp32-ops", "-mllvm", "-amdgpu-mfma-vgpr-form=1"] + SAN +
                    ["-c", os.path.join(CSRC, s), "-o", os.path.join(OUT, s[:-4] + ".o")])
    jobs.append([CLANG, "-O1", "-g", "-ffp-contract=off", "-fopenmp"] + HOST_SAN +
                ["-c", os.path.join(ROOT, "oracle", "oracle.c"), "-o", os.path.join(OUT, "oracle.o")])
    jobs.append([CLANG, "-O1", "-g", "-ffp-contract=off"] + HOST_SAN +
                ["-c", os.path.join(CSRC, "datagen.c"), "-o", os.path.join(OUT, "datagen.o")])
    jobs.append([CLANG + "++", "-O1", "-g", "-std=c++17"] + HOST_SAN +
                ["-c", os.path.join(ROOT, "tests", "sanitize_driver.cpp"), "-o", os.path.join(OUT, "driver.o")])
    with ThreadPoolExecutor(max_workers=8) as ex:
        list(ex.map(run, jobs))
    objs = [os.path.join(OUT, f) for f in sorted(os.listdir(OUT)) if f.endswith(".o")]
    exe = os.path.join(OUT, "sanitize_driver")
    run([HIPCC, "--offload-arch=gfx950", "-fopenmp"] + SAN + objs + ["-o", exe, "-lm"])
    return exe


def test_host_code_under_asan_and_ubsan(tmp_path):
    exe = build()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, str(tmp_path)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-6000:]
    assert "sanitize_driver ok" in r.stdout
