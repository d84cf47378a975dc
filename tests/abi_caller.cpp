// abi_caller.cpp — a C++ caller of libsift2nn.so written the way the
// reference's driver calls the reference (src/test.cu:45-215): descriptors as
// des_t_f arrays in page-locked mapped host memory and on the device, a
// float4 array for the 2-NN output, and the reference's calls with the
// reference's arguments.  What differs from src/test.cu is only what the
// platform changes: hip* allocation calls instead of cuda*, and the opaque
// sift_handle_t (sift_create) where the reference has a cublasHandle_t
// (cublasCreate).  Compiled by the tests against include/sift2nn.h alone.
//
//   abi_caller q.f32 r.f32 size_q size_r seed out_dir
//
// Inputs are raw float32 [n][128]; outputs (raw, in out_dir):
//   device_brute.f4   float4[size_q]   device_brute            (test.cu:160)
//   cublas_2nn_f.f4   float4[size_q]   cublas_2nn_f
//   sift_type0.u32    uint32[size_q]   cublas_2nn_sift type 0, threshold 2 (test.cu:210)
//   sift_type1.u32    uint32[size_q]   cublas_2nn_sift type 1, threshold 0.8
//   lsh.u32           uint32[size_q]   lsh_gpu type 1, 4 tables of 12 bits, threshold 0.8
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>

#include "sift2nn.h"

static bool read_all(const char* path, void* dst, size_t bytes) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    const bool ok = std::fread(dst, 1, bytes, f) == bytes;
    std::fclose(f);
    return ok;
}

static bool write_all(const std::string& path, const void* src, size_t bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(src, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}

#define CHECK(x)                                                              \
    do {                                                                      \
        if (!(x)) {                                                           \
            std::fprintf(stderr, "abi_caller: %s failed (line %d): %s\n", #x, \
                         __LINE__, sift_last_error(nullptr));                 \
            return 1;                                                         \
        }                                                                     \
    } while (0)

int main(int argc, char** argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: abi_caller q.f32 r.f32 size_q size_r seed out_dir\n");
        return 2;
    }
    const int size_q = std::atoi(argv[3]), size_r = std::atoi(argv[4]);
    const unsigned long long seed = std::strtoull(argv[5], nullptr, 10);
    const std::string out = argv[6];

    des_t_f* q_points;
    des_t_f* r_points;
    des_t_f* gpu_q_points;
    des_t_f* gpu_r_points;

    sift_handle_t handle;  // src/test.cu:51-52 creates a cuBLAS handle here
    CHECK(sift_create(&handle) == SIFT_OK);
    CHECK(sift_set_seed(handle, seed) == SIFT_OK);

    uint32_t* matches;
    CHECK(hipMalloc((void**)&matches, size_q * sizeof(uint32_t)) == hipSuccess);
    CHECK(hipHostMalloc((void**)&q_points, size_q * sizeof(des_t_f), hipHostMallocMapped) == hipSuccess);
    CHECK(hipHostMalloc((void**)&r_points, size_r * sizeof(des_t_f), hipHostMallocMapped) == hipSuccess);
    CHECK(hipMalloc((void**)&gpu_q_points, size_q * sizeof(des_t_f)) == hipSuccess);
    CHECK(hipMalloc((void**)&gpu_r_points, size_r * sizeof(des_t_f)) == hipSuccess);
    CHECK(read_all(argv[1], q_points, size_q * sizeof(des_t_f)));
    CHECK(read_all(argv[2], r_points, size_r * sizeof(des_t_f)));

    float4* sorted_2nn;
    CHECK(hipMalloc((void**)&sorted_2nn, size_q * sizeof(float4)) == hipSuccess);
    CHECK(hipMemcpy(gpu_q_points, q_points, size_q * sizeof(des_t_f), hipMemcpyHostToDevice) == hipSuccess);
    CHECK(hipMemcpy(gpu_r_points, r_points, size_r * sizeof(des_t_f), hipMemcpyHostToDevice) == hipSuccess);
    CHECK(hipDeviceSynchronize() == hipSuccess);

    float4* host4 = new float4[size_q];
    uint32_t* host_m = new uint32_t[size_q];

    device_brute(gpu_q_points, gpu_r_points, size_q, size_r, sorted_2nn);
    CHECK(sift_last_status() == SIFT_OK);
    CHECK(hipDeviceSynchronize() == hipSuccess);
    CHECK(hipMemcpy(host4, sorted_2nn, size_q * sizeof(float4), hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(write_all(out + "/device_brute.f4", host4, size_q * sizeof(float4)));

    CHECK(hipMemset(sorted_2nn, 0, size_q * sizeof(float4)) == hipSuccess);
    cublas_2nn_f(gpu_q_points, gpu_r_points, size_q, size_r, sorted_2nn, handle);
    CHECK(sift_last_status() == SIFT_OK);
    CHECK(hipMemcpy(host4, sorted_2nn, size_q * sizeof(float4), hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(write_all(out + "/cublas_2nn_f.f4", host4, size_q * sizeof(float4)));

    // the reference's timed call, src/test.cu:210
    CHECK(cublas_2nn_sift(q_points, r_points, 0, size_q, size_r, matches, 2, handle, 1,
                          (size_r * 2) * (size_t)383, 17) == SIFT_OK);
    CHECK(hipDeviceSynchronize() == hipSuccess);
    CHECK(hipMemcpy(host_m, matches, size_q * sizeof(uint32_t), hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(write_all(out + "/sift_type0.u32", host_m, size_q * sizeof(uint32_t)));

    CHECK(cublas_2nn_sift(gpu_q_points, gpu_r_points, 1, size_q, size_r, matches, 0.8f, handle, 1,
                          (size_r * 2) * (size_t)500, 17) == SIFT_OK);
    CHECK(hipMemcpy(host_m, matches, size_q * sizeof(uint32_t), hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(write_all(out + "/sift_type1.u32", host_m, size_q * sizeof(uint32_t)));

    // as the commented-out calls of src/test.cu:183 (l = 4, nbits = 12 here)
    CHECK(lsh_gpu(gpu_q_points, gpu_r_points, 1, size_q, size_r, matches, 0.8f, handle, 12, 4, 0, 12) == SIFT_OK);
    CHECK(hipMemcpy(host_m, matches, size_q * sizeof(uint32_t), hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(write_all(out + "/lsh.u32", host_m, size_q * sizeof(uint32_t)));

    delete[] host4;
    delete[] host_m;
    CHECK(hipFree(sorted_2nn) == hipSuccess);
    CHECK(hipFree(gpu_q_points) == hipSuccess);
    CHECK(hipFree(gpu_r_points) == hipSuccess);
    CHECK(hipHostFree(q_points) == hipSuccess);
    CHECK(hipHostFree(r_points) == hipSuccess);
    CHECK(hipFree(matches) == hipSuccess);
    CHECK(sift_destroy(handle) == SIFT_OK);
    std::printf("abi_caller ok\n");
    return 0;
}
