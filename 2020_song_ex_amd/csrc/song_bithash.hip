// song_bithash.hip -- BitHash projection on the GPU (the Hamming query path).
//
// Code bit B of a vector x (bit B % 64 of uint64 word B / 64) is
// sign(R[B] . x) > 0 for the 256 x dim Gaussian matrix R, as BitHash::project
// (bitHash.h:32-44) and song_bithash_project compute it: the dot product is
// summed over d = 0 .. dim-1 in order, in float, each product rounded before
// it is added.  The sums stay sequential on purpose: a tree or MFMA would
// reorder them and could flip the sign of a sum near zero, and the codes must
// equal the reference's bit for bit.
//
// One wave projects BH_VECS vectors.  Lane l owns rows l, l + 64, l + 128 and
// l + 192 of R, so the four ballots of one vector's signs are its four code
// words.  The vectors sit in LDS and are read as broadcasts; each lane walks
// its rows of R in order through L1 / L2 (R is shared by every wave).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "song_kernels.h"

namespace song {

namespace {

constexpr int BH_VECS = 8;  // vectors per wave

__global__ void __launch_bounds__(64) bithash_kernel(const float* __restrict__ R, int dim,
                                                     const float* __restrict__ X, int64_t n,
                                                     uint64_t* __restrict__ codes) {
    extern __shared__ float xs[];  // [BH_VECS][dim]
    const int lane = threadIdx.x;
    const int64_t v0 = (int64_t)blockIdx.x * BH_VECS;
    const int nv = (int)(n - v0 < BH_VECS ? n - v0 : BH_VECS);
    for (int i = lane; i < BH_VECS * dim; i += 64) xs[i] = i < nv * dim ? X[v0 * dim + i] : 0.f;
    __syncthreads();

    float acc[4][BH_VECS];
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int v = 0; v < BH_VECS; v++) acc[j][v] = 0.f;
    const float* r = R + (size_t)lane * dim;
    const size_t quarter = (size_t)64 * dim;
    for (int d = 0; d < dim; d++) {
        float rj[4];
#pragma unroll
        for (int j = 0; j < 4; j++) rj[j] = r[j * quarter + d];
#pragma unroll
        for (int v = 0; v < BH_VECS; v++) {
            const float x = xs[v * dim + d];
#pragma unroll
            for (int j = 0; j < 4; j++) acc[j][v] = __fadd_rn(acc[j][v], __fmul_rn(rj[j], x));
        }
    }
#pragma unroll
    for (int v = 0; v < BH_VECS; v++) {
        if (v < nv) {
            uint64_t w = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint64_t b = __ballot(acc[j][v] > 0.f);
                if (lane == j) w = b;
            }
            if (lane < 4) codes[(v0 + v) * 4 + lane] = w;
        }
    }
}

}  // namespace

hipError_t launch_bithash(const float* R, int dim, const float* X, int64_t n, uint64_t* codes, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    const int64_t blocks = (n + BH_VECS - 1) / BH_VECS;
    const size_t lds = (size_t)BH_VECS * dim * sizeof(float);
    hipLaunchKernelGGL(bithash_kernel, dim3((unsigned)blocks), dim3(64), lds, st, R, dim, X, n, codes);
    return hipGetLastError();
}

}  // namespace song
