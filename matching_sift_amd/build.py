"""Build the sift2nn shared libraries in-tree (gfx950 only).

libsift2nn.so  : HIP kernels + host C++ behind the C ABI of include/sift2nn.h
libsiftdata.so : seeded synthetic-descriptor generator (host C)

Every call compiles from scratch; objects go to matching_sift_amd/_build/.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libsift2nn.so")
DATALIB = os.path.join(PKG, "libsiftdata.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-Wall",
    "-Wno-unused-result",
    # packed FP32 ops are kept out of generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

# fused.hip: keys are never NaN, so min() needs no NaN canonicalisation
FILE_FLAGS = {"fused.hip": ["-fno-honor-nans", "-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    if verbose and out:
        print(out)
    return out


def build(verbose=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(BUILD, s[:-4] + ".o") for s in srcs]

    def compile_one(i):
        extra = FILE_FLAGS.get(srcs[i], [])
        return _run([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, srcs[i]), "-o", objs[i]],
                    verbose)

    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        list(ex.map(compile_one, range(len(srcs))))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    tmp = DATALIB + ".tmp"
    _run(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", tmp,
          os.path.join(CSRC, "datagen.c"), "-lm"], verbose)
    os.replace(tmp, DATALIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB, DATALIB)
