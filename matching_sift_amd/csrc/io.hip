// io.hip — SIFT vector files and the recall count against a ground truth.
//
// File formats of the ANN_SIFT1M set as the reference reads them
// (src/test.cu:76-155): every record is an int32 dimension d followed by d
// elements — float32 (.fvecs), int32 (.ivecs, ground-truth neighbour ids) or
// uint8 (.bvecs).  The reference reads 128-dim fvecs rows (optionally / 512)
// and the first `recall` of the 100 ground-truth ids per query.
//
// The recall count follows the reference's loop (src/test.cu:224-301): a
// query counts as found at depth R when its returned match is among the first
// R ground-truth ids.  Host code only.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"

using namespace sift;

namespace {

size_t elem_size(int elem) { return elem == SIFT_VECS_U8 ? 1 : 4; }

struct File {
    FILE* f = nullptr;
    explicit File(const char* path) : f(path ? std::fopen(path, "rb") : nullptr) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

// Record layout of an open file: dimension from the first header, count from
// the file size (which must be a whole number of records).
int layout(FILE* f, const char* path, int elem, uint32_t* n, uint32_t* dim) {
    int32_t d = 0;
    if (std::fseek(f, 0, SEEK_END) != 0) return set_error(nullptr, SIFT_ERR_INVALID, std::string("seek ") + path);
    const long size = std::ftell(f);
    if (size < 0) return set_error(nullptr, SIFT_ERR_INVALID, std::string("size of ") + path);
    if (size == 0) {
        *n = 0;
        *dim = 0;
        return SIFT_OK;
    }
    std::rewind(f);
    if (std::fread(&d, 4, 1, f) != 1 || d <= 0 || d > (1 << 20))
        return set_error(nullptr, SIFT_ERR_INVALID, std::string("bad dimension header in ") + path);
    const size_t rec = 4 + (size_t)d * elem_size(elem);
    if ((size_t)size % rec != 0)
        return set_error(nullptr, SIFT_ERR_INVALID, std::string(path) + ": size is not a whole number of records");
    if ((size_t)size / rec > 0xFFFFFFFFull) return set_error(nullptr, SIFT_ERR_UNSUPPORTED, "too many records");
    *n = (uint32_t)((size_t)size / rec);
    *dim = (uint32_t)d;
    return SIFT_OK;
}

}  // namespace

extern "C" {

int sift_vecs_info(const char* path, int elem, uint32_t* n, uint32_t* dim) {
    if (!path || !n || !dim) return set_error(nullptr, SIFT_ERR_INVALID, "null argument");
    if (elem < SIFT_VECS_F32 || elem > SIFT_VECS_U8) return set_error(nullptr, SIFT_ERR_INVALID, "unknown element kind");
    File f(path);
    if (!f.f) return set_error(nullptr, SIFT_ERR_INVALID, std::string("cannot open ") + path);
    return layout(f.f, path, elem, n, dim);
}

int sift_vecs_read(const char* path, int elem, uint64_t first, uint32_t n, uint32_t cols, float scale, void* out) {
    if (!path || (n && !out)) return set_error(nullptr, SIFT_ERR_INVALID, "null argument");
    if (elem < SIFT_VECS_F32 || elem > SIFT_VECS_U8) return set_error(nullptr, SIFT_ERR_INVALID, "unknown element kind");
    File f(path);
    if (!f.f) return set_error(nullptr, SIFT_ERR_INVALID, std::string("cannot open ") + path);
    uint32_t total = 0, d = 0;
    SIFT_TRY(layout(f.f, path, elem, &total, &d));
    if (first + n > total) return set_error(nullptr, SIFT_ERR_INVALID, std::string(path) + ": fewer records than asked");
    if (cols == 0) cols = d;
    if (cols > d) return set_error(nullptr, SIFT_ERR_INVALID, "cols exceeds the file's dimension");
    const size_t es = elem_size(elem), rec = 4 + (size_t)d * es;
    if (std::fseek(f.f, (long)(first * rec), SEEK_SET) != 0) return set_error(nullptr, SIFT_ERR_INVALID, "seek");
    std::vector<unsigned char> buf(rec);
    for (uint32_t i = 0; i < n; i++) {
        if (std::fread(buf.data(), 1, rec, f.f) != rec) return set_error(nullptr, SIFT_ERR_INVALID, "short read");
        int32_t h = 0;
        std::memcpy(&h, buf.data(), 4);
        if ((uint32_t)h != d)
            return set_error(nullptr, SIFT_ERR_INVALID, std::string(path) + ": record " + std::to_string(first + i) +
                                                            " has another dimension");
        const unsigned char* p = buf.data() + 4;
        if (elem == SIFT_VECS_I32) {
            std::memcpy(static_cast<int32_t*>(out) + (size_t)i * cols, p, (size_t)cols * 4);
        } else {
            float* o = static_cast<float*>(out) + (size_t)i * cols;
            for (uint32_t c = 0; c < cols; c++) {
                float v;
                if (elem == SIFT_VECS_F32)
                    std::memcpy(&v, p + 4 * (size_t)c, 4);
                else
                    v = (float)p[c];
                o[c] = v * scale;
            }
        }
    }
    return SIFT_OK;
}

int sift_recall_count(const uint32_t* found, uint32_t q_n, const int32_t* gt, uint32_t gt_k,
                      const uint32_t* depths, int n_depths, uint32_t* hits) {
    if (n_depths < 0 || (n_depths && (!depths || !hits)) || (q_n && (!found || !gt)))
        return set_error(nullptr, SIFT_ERR_INVALID, "null argument");
    for (int k = 0; k < n_depths; k++) {
        if (depths[k] < 1 || depths[k] > gt_k) return set_error(nullptr, SIFT_ERR_INVALID, "depth outside [1, gt_k]");
        hits[k] = 0;
    }
    for (uint32_t q = 0; q < q_n; q++) {
        // position of the returned match in the query's ground-truth list
        uint32_t pos = gt_k;
        if (found[q] != 0xFFFFFFFFu)
            for (uint32_t j = 0; j < gt_k; j++)
                if ((uint32_t)gt[(size_t)q * gt_k + j] == found[q]) {
                    pos = j;
                    break;
                }
        for (int k = 0; k < n_depths; k++) hits[k] += pos < depths[k] ? 1u : 0u;
    }
    return SIFT_OK;
}

}  // extern "C"
