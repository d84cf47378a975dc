// lsh.hip — random-hyperplane LSH index.
#include "internal.h"

extern "C" void sift_lsh_state_free(sift_context* h) { (void)h; }

extern "C" int lsh_gpu(void* q_points, void* r_points, int type, uint32_t q_n, uint32_t r_n,
                       uint32_t* matches, float threshold, sift_handle_t handle, int stream_n,
                       int l, int lsh_type, int nbits) {
    (void)q_points; (void)r_points; (void)type; (void)q_n; (void)r_n; (void)matches;
    (void)threshold; (void)stream_n; (void)l; (void)lsh_type; (void)nbits;
    return sift::set_error(handle, SIFT_ERR_UNSUPPORTED, "lsh_gpu: not built yet");
}
