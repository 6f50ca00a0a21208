// bithash.cpp -- 1-bit Gaussian random projection to 256-bit codes.
//
// Same matrix and bits as the reference's BitHash (bitHash.h:17-56):
// R is [256 x dim], filled row-major from std::mt19937(seed) through
// std::normal_distribution<float>(0, 1); bit b of 32-bit word w is
// sign(R[32w + b] . x) > 0, the dot product summed sequentially in float
// with each product rounded on its own.  Codes are returned as 4 x uint64
// (32-bit word w sits at bit 32*(w%2) of uint64 word w/2).
#include <cstdint>
#include <random>
#include <thread>
#include <vector>

extern "C" {

int song_bithash_matrix(int32_t dim, uint32_t seed, float* out) {
    if (dim < 1 || !out) return -1;
    std::mt19937 rng(seed);
    std::normal_distribution<float> gauss(0.0f, 1.0f);
    for (int64_t i = 0; i < (int64_t)256 * dim; i++) out[i] = gauss(rng);
    return 0;
}

int song_bithash_project(const float* R, int32_t dim, const float* vecs, int64_t n, int32_t n_threads,
                         uint64_t* codes_out) {
    if (dim < 1 || !R || (n > 0 && (!vecs || !codes_out)) || n < 0) return -1;
    auto work = [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++) {
            const float* x = vecs + i * dim;
            uint64_t* c = codes_out + i * 4;
            for (int w = 0; w < 8; w++) {
                uint32_t word = 0;
                for (int b = 0; b < 32; b++) {
                    const float* r = R + (int64_t)(w * 32 + b) * dim;
                    float dot = 0.f;
                    for (int d = 0; d < dim; d++) {
                        float p = r[d] * x[d];
                        dot = dot + p;
                    }
                    if (dot > 0) word |= (1u << b);
                }
                if (w % 2 == 0) c[w / 2] = word;
                else c[w / 2] |= (uint64_t)word << 32;
            }
        }
    };
    if (n_threads <= 1 || n < 1024) {
        work(0, n);
    } else {
        std::vector<std::thread> th;
        int64_t per = (n + n_threads - 1) / n_threads;
        for (int t = 0; t < n_threads; t++) {
            int64_t lo = t * per, hi = std::min<int64_t>(n, lo + per);
            if (lo < hi) th.emplace_back(work, lo, hi);
        }
        for (auto& x : th) x.join();
    }
    return 0;
}

}  // extern "C"
