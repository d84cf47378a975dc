// sanitize_driver.cpp — exercises the host-only code paths under AddressSanitizer
// and UndefinedBehaviorSanitizer on the CPU (no device call is made):
//   - the SIFT vector file reader and the recall count (io.hip), on valid,
//     damaged and byte-flipped files;
//   - the LSH index file checker (sift_lsh_check_file), on a valid file and
//     on truncated and byte-flipped copies of it;
//   - the CPU oracle (oracle/oracle.c) and the synthetic data generator
//     (csrc/datagen.c) on small inputs, edge sizes included.
// Every call must return normally (an error code for damaged input); the
// sanitizers abort the process on any invalid access or undefined behaviour.
//
//   sanitize_driver <scratch dir>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/sift2nn.h"
#include "../oracle/oracle.h"

extern "C" {
void sift_gen_base(uint64_t seed, int64_t n, float* out);
void sift_gen_queries(uint64_t seed, const float* base, int64_t n_base, int64_t m, double frac_copy,
                      double noise_sd, float* out, int64_t* src);
void sift_gen_pair_sizes(uint64_t seed, int64_t pairs, double median, double sigma, int64_t lo, int64_t hi,
                         int64_t* out);
void sift_gen_unit(uint64_t seed, int64_t n, float* out);
void sift_gen_gauss_int(uint64_t seed, int64_t n, double sd, double lim, float* out);
void sift_gen_uniform(uint64_t seed, int64_t n, double* out);
}

namespace {

uint64_t g_state = 0x243F6A8885A308D3ull;
uint32_t rnd() {  // xorshift64*, deterministic
    g_state ^= g_state >> 12;
    g_state ^= g_state << 25;
    g_state ^= g_state >> 27;
    return (uint32_t)((g_state * 0x2545F4914F6CDD1Dull) >> 32);
}

void write_file(const std::string& path, const std::vector<uint8_t>& bytes) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    if (!bytes.empty()) std::fwrite(bytes.data(), 1, bytes.size(), f);
    std::fclose(f);
}

template <typename T>
void put(std::vector<uint8_t>& b, T v) {
    const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
    b.insert(b.end(), p, p + sizeof(T));
}

// n records of dimension d, element size es (4: float/int32, 1: uint8)
std::vector<uint8_t> vecs_file(uint32_t n, int32_t d, int es) {
    std::vector<uint8_t> b;
    for (uint32_t i = 0; i < n; i++) {
        put<int32_t>(b, d);
        for (int32_t k = 0; k < d; k++) {
            if (es == 1) b.push_back((uint8_t)rnd());
            else put<int32_t>(b, (int32_t)(rnd() % 1000));
        }
    }
    return b;
}

int checks = 0;
void vecs_calls(const std::string& path, int elem) {
    uint32_t n = 0, dim = 0;
    const int rc = sift_vecs_info(path.c_str(), elem, &n, &dim);
    checks++;
    if (rc != SIFT_OK || n == 0 || dim == 0 || dim > 4096) return;
    const uint32_t take = n < 8 ? n : 8;
    for (uint32_t cols : {0u, 1u, dim}) {
        std::vector<float> out((size_t)take * (cols ? cols : dim) + 1);
        sift_vecs_read(path.c_str(), elem, 0, take, cols, 0.5f, out.data());
        sift_vecs_read(path.c_str(), elem, n - take, take, cols, 1.f, out.data());
        sift_vecs_read(path.c_str(), elem, n, 1, cols, 1.f, out.data());  // past the end: refused
        checks += 3;
    }
}

void test_vecs(const std::string& dir) {
    const std::string p = dir + "/x.vecs";
    for (int es : {4, 1}) {
        const int elem = es == 1 ? SIFT_VECS_U8 : SIFT_VECS_F32;
        const std::vector<uint8_t> good = vecs_file(13, 128, es);
        write_file(p, good);
        vecs_calls(p, elem);
        vecs_calls(p, SIFT_VECS_I32);
        for (size_t cut : {(size_t)0, (size_t)1, (size_t)3, (size_t)4, good.size() - 1}) {
            write_file(p, std::vector<uint8_t>(good.begin(), good.begin() + (long)cut));
            vecs_calls(p, elem);
        }
        for (int it = 0; it < 200; it++) {  // byte flips, headers included
            std::vector<uint8_t> bad = good;
            const int flips = 1 + (int)(rnd() % 4);
            for (int k = 0; k < flips; k++) {
                const size_t at = (rnd() % 3 == 0) ? (size_t)(rnd() % 4) : rnd() % bad.size();
                bad[at] ^= (uint8_t)(1u << (rnd() % 8));
            }
            write_file(p, bad);
            vecs_calls(p, elem);
        }
    }
    std::vector<uint8_t> huge;
    put<int32_t>(huge, 0x7FFFFFFF);
    write_file(p, huge);
    vecs_calls(p, SIFT_VECS_F32);
    put<int32_t>(huge, -5);
    write_file(p, huge);
    vecs_calls(p, SIFT_VECS_F32);
    vecs_calls(dir + "/missing.vecs", SIFT_VECS_F32);
    // recall count: ids in and out of range, UINT32_MAX, depths at the limits
    const uint32_t qn = 9, gk = 7;
    std::vector<uint32_t> found(qn);
    std::vector<int32_t> gt((size_t)qn * gk);
    for (auto& v : gt) v = (int32_t)(rnd() % 20) - 2;
    for (uint32_t i = 0; i < qn; i++) found[i] = i % 3 == 0 ? 0xFFFFFFFFu : rnd() % 20;
    std::vector<uint32_t> depths = {1, 3, gk}, hits(3);
    sift_recall_count(found.data(), qn, gt.data(), gk, depths.data(), 3, hits.data());
    depths = {0, gk + 1, 2};
    sift_recall_count(found.data(), qn, gt.data(), gk, depths.data(), 3, hits.data());  // refused
    sift_recall_count(found.data(), 0, gt.data(), gk, depths.data(), 1, hits.data());
    checks += 3;
}

// The version-1 LSH index layout (see lsh.hip): header, hyperplanes, keys,
// rows, uint8 rows, int32 norms.
std::vector<uint8_t> lsh_file(uint32_t n, int32_t nbits, int32_t L) {
    std::vector<uint8_t> b;
    const char magic[8] = {'S', 'I', 'F', 'T', 'L', 'S', 'H', '1'};
    b.insert(b.end(), magic, magic + 8);
    put<uint32_t>(b, 1);
    put<uint32_t>(b, n);
    put<int32_t>(b, nbits);
    put<int32_t>(b, L);
    put<int32_t>(b, SIFT_LSH_HYPERPLANE);
    put<uint32_t>(b, 0);
    put<uint64_t>(b, 7);
    for (int32_t i = 0; i < L * nbits * 128; i++) put<float>(b, (float)((int)(rnd() % 255) - 127));
    std::vector<uint32_t> keys, vals;
    for (int32_t t = 0; t < L; t++)
        for (uint32_t j = 0; j < n; j++) {
            keys.push_back(((uint32_t)t << nbits) | (j * 3 / (n ? n : 1)));
            vals.push_back(j);
        }
    for (uint32_t k : keys) put<uint32_t>(b, k);
    for (uint32_t v : vals) put<uint32_t>(b, v);
    std::vector<uint8_t> rows((size_t)n * 128);
    for (auto& x : rows) x = (uint8_t)rnd();
    b.insert(b.end(), rows.begin(), rows.end());
    for (uint32_t j = 0; j < n; j++) {
        int32_t s = 0;
        for (int d = 0; d < 128; d++) s += (int32_t)rows[(size_t)j * 128 + d] * rows[(size_t)j * 128 + d];
        put<int32_t>(b, s);
    }
    return b;
}

void test_lsh_files(const std::string& dir) {
    const std::string p = dir + "/x.lsh";
    const std::vector<uint8_t> good = lsh_file(9, 4, 3);
    write_file(p, good);
    if (sift_lsh_check_file(p.c_str()) != SIFT_OK) {
        std::fprintf(stderr, "the well-formed index file was refused: %s\n", sift_last_error(nullptr));
        std::exit(1);
    }
    for (size_t cut : {(size_t)0, (size_t)7, (size_t)40, good.size() / 2, good.size() - 1}) {
        write_file(p, std::vector<uint8_t>(good.begin(), good.begin() + (long)cut));
        if (sift_lsh_check_file(p.c_str()) == SIFT_OK) {
            std::fprintf(stderr, "a truncated index file was accepted\n");
            std::exit(1);
        }
    }
    for (int it = 0; it < 400; it++) {
        std::vector<uint8_t> bad = good;
        const int flips = 1 + (int)(rnd() % 3);
        for (int k = 0; k < flips; k++) {
            const size_t at = (rnd() % 2 == 0) ? (size_t)(rnd() % 40) : rnd() % bad.size();
            bad[at] ^= (uint8_t)(1u << (rnd() % 8));
        }
        write_file(p, bad);
        sift_lsh_check_file(p.c_str());
        checks++;
    }
    sift_lsh_check_file((dir + "/missing.lsh").c_str());
}

void test_oracle_and_datagen() {
    for (int64_t n : {0, 1, 2, 33}) {
        std::vector<float> b((size_t)n * 128 + 1), q(5 * 128);
        std::vector<int64_t> src(5);
        sift_gen_base(3, n, b.data());
        if (n > 0) sift_gen_queries(4, b.data(), n, 5, 0.5, 8.0, q.data(), src.data());
        else sift_gen_base(5, 5, q.data());
        std::vector<float> s4(5 * 4), dm((size_t)5 * (n ? n : 1));
        oracle_host_brute(q.data(), b.data(), 5, (int)n, s4.data());
        oracle_host_brute_mt(q.data(), b.data(), 5, (int)n, s4.data(), 2);
        for (int m = 1; m <= 4; m++) oracle_host_brute_mutant(q.data(), b.data(), 5, (int)n, s4.data(), m);
        std::vector<uint32_t> idx2(10), dist2(10), matches(5), nc(5);
        oracle_knn2_int(q.data(), b.data(), 5, (int)n, idx2.data(), dist2.data(), 2);
        oracle_ratio(s4.data(), 5, 0.8f, matches.data());
        if (n > 0) {
            for (size_t i = 0; i < dm.size(); i++) dm[i] = (float)(rnd() % 100);
            oracle_host_sort(dm.data(), (int)n, 5, s4.data());
            std::vector<int64_t> qi = {0, 4}, ri = {0, n - 1};
            std::vector<float> d2(2);
            oracle_host_lenght(q.data(), b.data(), qi.data(), ri.data(), 2, d2.data());
        }
        std::vector<float> H(6 * 128);
        for (auto& h : H) h = (float)((int)(rnd() % 255) - 127);
        std::vector<uint32_t> cq(2 * 5), cr(2 * (size_t)(n ? n : 1)), perm((size_t)(n ? n : 1));
        oracle_lsh_codes(q.data(), 5, H.data(), 6, cq.data());
        oracle_lsh_codes(q.data(), 5, H.data() + 3 * 128, 3, cq.data() + 5);
        oracle_lsh_codes(b.data(), (int)n, H.data(), 6, cr.data());
        oracle_lsh_codes(b.data(), (int)n, H.data() + 3 * 128, 3, cr.data() + n);
        oracle_lsh_sort(cr.data(), (int)n, perm.data());
        oracle_lsh_query(q.data(), b.data(), 5, (int)n, 2, cq.data(), cr.data(), idx2.data(), dist2.data(), nc.data(), 2);
        std::vector<float> fd2(10), qq(5 * 128);
        oracle_lsh_query_f32(q.data(), b.data(), 5, (int)n, 2, cq.data(), cr.data(), idx2.data(), fd2.data(), nc.data(), 2);
        const int k = oracle_lsh_scale_exp(q.data(), 5 * 128);
        oracle_lsh_quantize(q.data(), 5 * 128, k, qq.data());
        checks += 12;
    }
    std::vector<float> u(3 * 128), g(50);
    std::vector<double> un(17);
    std::vector<int64_t> ps(6 * 2);
    sift_gen_unit(9, 3, u.data());
    sift_gen_gauss_int(10, 50, 32.0, 127.0, g.data());
    sift_gen_uniform(11, 17, un.data());
    sift_gen_pair_sizes(12, 6, 2000.0, 0.6, 200, 8000, ps.data());
    // learned-hash oracle on 2 rows
    const int nb = 12, NP = SIFT_MLP_NPARAM;
    std::vector<float> params(NP);
    for (auto& v : params) v = (float)((int)(rnd() % 200) - 100) / 100.f;
    std::vector<double> hid(2 * 256), z(2 * 32), dls(2 * 32), dz(2 * 32), grads(NP), out(2 * 8), hid2(2 * 16);
    oracle_hash_forward(params.data(), nb, u.data(), 2, 1, hid.data(), z.data());
    oracle_hash_forward(params.data(), nb, u.data(), 2, 0, hid.data(), z.data());
    std::vector<double> sgn(2 * 32);
    for (size_t i = 0; i < sgn.size(); i++) sgn[i] = z[i] < 0 ? -1.0 : 1.0;
    oracle_hash_loss(sgn.data(), z.data(), 1, nb, 1.0, 1.0, 0.1);
    oracle_hash_loss_grad(z.data(), 1, nb, 1.0, 1.0, 0.1, dls.data(), dz.data());
    oracle_hash_backward(params.data(), nb, u.data(), 2, 1, hid.data(), dz.data(), grads.data());
    oracle_mlp_forward(u.data(), 2, params.data(), params.data() + 256 * 128, 16, params.data(), params.data(), 8,
                       hid2.data(), out.data());
    checks += 8;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: sanitize_driver <scratch dir>\n");
        return 2;
    }
    const std::string dir = argv[1];
    test_vecs(dir);
    test_lsh_files(dir);
    test_oracle_and_datagen();
    std::printf("sanitize_driver ok: %d checks\n", checks);
    return 0;
}
