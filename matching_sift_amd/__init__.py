"""matching_sift_amd — MI355X (gfx950) 2-NN SIFT matching.

The product is libsift2nn.so (HIP kernels + host C++ behind include/sift2nn.h).
This package loads it with ctypes and mirrors the reference's entry points
(src/knn_brute.cu, src/lsh_h.cu) for Python callers, tests and the bench.
There is no CPU fallback: if the library is missing or no GPU is present the
calls raise.
"""
from ._lib import (  # noqa: F401
    LIB_PATH, SiftError, lib, load,
    FLOAT_HOST, FLOAT_DEVICE, HALF_DEVICE, BF16_DEVICE, U8_DEVICE,
    COMPUTE_AUTO, COMPUTE_BF16, COMPUTE_F16, COMPUTE_F32, COMPUTE_I8, COMPUTE_FP8,
    LSH_HYPERPLANE, LSH_LEARNED, LSH_HYPERPLANE_PM1,
)
from .api import (  # noqa: F401
    DeviceArray, PinnedArray, Handle, to_device, knn2, cublas_2nn_sift, cublas_2nn_sift_batch,
    device_brute, cublas_2nn_f, cublas_2nn_brute_f, lsh_check_file,
    BruteForceIndex, LSHIndex, lsh_gpu, PairsPlan, match_pairs, PAIRS_GRAPH, PAIRS_AUTOTUNE,
    HashMLP, TrainOpts, MLP_NPARAM, vecs_info, read_vecs, recall_count,
)
