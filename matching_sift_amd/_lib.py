"""ctypes binding of libsift2nn.so (declarations follow include/sift2nn.h)."""
import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libsift2nn.so")

FLOAT_HOST, FLOAT_DEVICE, HALF_DEVICE, BF16_DEVICE, U8_DEVICE = 0, 1, 2, 3, 4
COMPUTE_AUTO, COMPUTE_BF16, COMPUTE_F16, COMPUTE_F32, COMPUTE_I8, COMPUTE_FP8 = 0, 1, 2, 3, 4, 5
LSH_HYPERPLANE, LSH_LEARNED, LSH_HYPERPLANE_PM1 = 0, 1, 2

ERRORS = {-1: "invalid argument", -2: "out of device memory", -3: "HIP runtime error",
          -4: "unsupported", -5: "data outside the exact domain"}


class SiftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("sift2nn error %d (%s): %s" % (code, ERRORS.get(code, "?"), msg))
        self.code = code


_lib = None

vp, u32, u64, i32, f32, sz = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_float, C.c_size_t

# name -> (restype, argtypes); every function of include/sift2nn.h
SIGNATURES = {
    "sift_create": (i32, [C.POINTER(vp)]),
    "sift_vecs_info": (i32, [C.c_char_p, i32, C.POINTER(u32), C.POINTER(u32)]),
    "sift_vecs_read": (i32, [C.c_char_p, i32, u64, u32, u32, f32, vp]),
    "sift_recall_count": (i32, [vp, u32, vp, u32, vp, i32, vp]),
    "sift_destroy": (i32, [vp]),
    "sift_last_error": (C.c_char_p, [vp]),
    "sift_last_status": (i32, []),
    "sift_synchronize": (i32, [vp]),
    "sift_get_stream": (vp, [vp]),
    "sift_device_name": (i32, [C.c_char_p, i32]),
    "sift_malloc": (i32, [C.POINTER(vp), sz]),
    "sift_free": (i32, [vp]),
    "sift_host_alloc": (i32, [C.POINTER(vp), sz]),
    "sift_host_free": (i32, [vp]),
    "sift_memcpy_h2d": (i32, [vp, vp, sz]),
    "sift_memcpy_d2h": (i32, [vp, vp, sz]),
    "sift_memset": (i32, [vp, i32, sz]),
    "sift_profile_enable": (i32, [vp, i32]),
    "sift_profile_read": (i32, [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_long)]),
    "sift_knn2_shape": (i32, [vp, i32, u32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
    "sift_profile_reset": (i32, [vp]),
    "cublas_2nn_sift": (i32, [vp, vp, i32, u32, u32, vp, f32, vp, i32, sz, i32]),
    "cublas_2nn_sift_batch": (i32, [vp, vp, i32, vp, u32, u32, vp, vp, f32, vp, vp, i32]),
    "device_brute": (None, [vp, vp, i32, i32, vp]),
    "cublas_2nn_f": (None, [vp, vp, i32, i32, vp, vp]),
    "cublas_2nn_brute_f": (None, [vp, vp, i32, i32, vp, vp, vp]),
    "lsh_gpu": (i32, [vp, vp, i32, u32, u32, vp, f32, vp, i32, i32, i32, i32]),
    "sift_set_seed": (i32, [vp, u64]),
    "sift_knn2": (i32, [vp, vp, vp, i32, u32, u32, i32, vp, vp, vp, vp, f32]),
    "sift_knn2_async": (i32, [vp, vp, vp, i32, u32, u32, i32, vp, vp, vp, vp, f32]),
    "sift_bf_index_build": (i32, [vp, vp, i32, u32, i32, C.POINTER(vp)]),
    "sift_bf_index_query": (i32, [vp, vp, vp, i32, u32, vp, vp, vp, vp, f32, i32]),
    "sift_bf_index_destroy": (i32, [vp]),
    "sift_bf_index_query_noalloc": (i32, [vp, vp, vp, i32, u32, vp, vp, vp, vp, f32, i32]),
    "sift_lsh_build": (i32, [vp, vp, i32, u32, i32, i32, i32, u64, C.POINTER(vp)]),
    "sift_lsh_query": (i32, [vp, vp, vp, i32, u32, vp, vp, vp, vp, f32, i32]),
    "sift_lsh_destroy": (i32, [vp]),
    "sift_lsh_codes": (i32, [vp, vp, vp, i32, u32, vp]),
    "sift_lsh_info": (i32, [vp, C.POINTER(u32), C.POINTER(i32), C.POINTER(i32)]),
    "sift_lsh_hash_input": (i32, [vp, C.POINTER(i32), C.POINTER(i32)]),
    "sift_lsh_get_hyperplanes": (i32, [vp, vp]),
    "sift_lsh_get_buckets": (i32, [vp, vp, vp]),
    "sift_mlp_create": (i32, [vp, i32, u64, C.POINTER(vp)]),
    "sift_mlp_destroy": (i32, [vp]),
    "sift_mlp_get_params": (i32, [vp, vp]),
    "sift_mlp_set_params": (i32, [vp, vp]),
    "sift_mlp_codes": (i32, [vp, vp, vp, i32, u32, vp, vp]),
    "sift_mlp_grad": (i32, [vp, vp, vp, vp, i32, u32, vp, vp, vp, vp]),
    "sift_mlp_train": (i32, [vp, vp, vp, u32, vp, u32, i32, vp, vp]),
    "sift_lsh_build_learned": (i32, [vp, vp, i32, u32, vp, i32, C.POINTER(vp)]),
    "sift_lsh_set_probes": (i32, [vp, i32]),
    "sift_lsh_bucket_stat": (i32, [vp, C.POINTER(C.c_double)]),
    "sift_lsh_probe_codes": (i32, [vp, vp, vp, i32, u32, vp]),
    "sift_lsh_save": (i32, [vp, C.c_char_p]),
    "sift_lsh_load": (i32, [vp, C.c_char_p, C.POINTER(vp)]),
    "sift_lsh_check_file": (i32, [C.c_char_p]),
    "sift_lsh_query_fallback": (i32, [vp, vp, vp, vp, i32, u32, u32, f32, vp, vp, vp, vp, f32, vp]),
    "sift_pairs_plan_create": (i32, [vp, i32, i32, u32, vp, u32, vp, u32, vp, C.c_uint, C.POINTER(vp)]),
    "sift_pairs_plan_execute": (i32, [vp, vp, vp, vp, vp, vp, vp, f32, i32]),
    "sift_pairs_plan_set_shape": (i32, [vp, i32, i32, i32]),
    "sift_pairs_plan_info": (i32, [vp, C.POINTER(C.c_int64)]),
    "sift_pairs_plan_destroy": (i32, [vp]),
    "sift_match_pairs": (i32, [vp, vp, vp, i32, u32, vp, u32, vp, u32, vp, i32, vp, vp, vp, f32]),
}


def load(path=LIB_PATH):
    """Load libsift2nn.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError("libsift2nn.so not built (%s); run __graft_entry__.build()" % path)
    L = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def lib():
    return load()


def check(rc, handle=None):
    if rc != 0:
        msg = lib().sift_last_error(handle)
        raise SiftError(rc, msg.decode() if msg else "")
    return rc
