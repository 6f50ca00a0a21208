// song_search.hip -- search-kernel dispatch by metric and set size (the
// kernels themselves are instantiated by the song_search_part objects), and
// the visit-table kernel.
#include <algorithm>

#include "song_search_impl.h"

namespace song {

#define SONG_PARTS_OF(M)                                                                                       \
    hipError_t launch_search_part_##M##_1(const SearchArgs&, int, int, const LaunchShape&, hipStream_t);       \
    hipError_t launch_search_part_##M##_2(const SearchArgs&, int, int, const LaunchShape&, hipStream_t);       \
    hipError_t launch_search_part_##M##_3(const SearchArgs&, int, int, const LaunchShape&, hipStream_t);       \
    hipError_t launch_search_part_##M##_4(const SearchArgs&, int, int, const LaunchShape&, hipStream_t);       \
    hipError_t launch_search_part_##M##_5(const SearchArgs&, int, int, const LaunchShape&, hipStream_t);       \
    hipError_t launch_search_part_##M##_6(const SearchArgs&, int, int, const LaunchShape&, hipStream_t);       \
    hipError_t launch_search_part_##M##_7(const SearchArgs&, int, int, const LaunchShape&, hipStream_t);       \
    hipError_t launch_search_part_##M##_8(const SearchArgs&, int, int, const LaunchShape&, hipStream_t);       \
    static hipError_t launch_metric_##M(const SearchArgs& a, int S, int G, int qch, const LaunchShape& ls,     \
                                        hipStream_t st) {                                                      \
        switch (S) {                                                                                           \
        case 1: return launch_search_part_##M##_1(a, G, qch, ls, st);                                          \
        case 2: return launch_search_part_##M##_2(a, G, qch, ls, st);                                          \
        case 3: return launch_search_part_##M##_3(a, G, qch, ls, st);                                          \
        case 4: return launch_search_part_##M##_4(a, G, qch, ls, st);                                          \
        case 5: return launch_search_part_##M##_5(a, G, qch, ls, st);                                          \
        case 6: return launch_search_part_##M##_6(a, G, qch, ls, st);                                          \
        case 7: return launch_search_part_##M##_7(a, G, qch, ls, st);                                          \
        case 8: return launch_search_part_##M##_8(a, G, qch, ls, st);                                          \
        }                                                                                                      \
        return hipErrorInvalidValue;                                                                           \
    }

// metric ids as in song_common.h: 0 L2, 1 inner product, 2 Hamming
SONG_PARTS_OF(0)
SONG_PARTS_OF(1)
SONG_PARTS_OF(2)

hipError_t launch_search(const SearchArgs& a, int metric, int S, int G, int qch, const LaunchShape& shape,
                         hipStream_t st) {
    switch (metric) {
    case METRIC_L2: return launch_metric_0(a, S, G, qch, shape, st);
    case METRIC_IP: return launch_metric_1(a, S, G, qch, shape, st);
    case METRIC_HAMMING: return launch_metric_2(a, S, G, qch, shape, st);
    }
    return hipErrorInvalidValue;
}

// One thread per adjacency slot: the low 16 bits of the slot id's four
// filter hashes, planes as SearchArgs::vpos lays them out.
static __global__ void visit_table_kernel(const int32_t* __restrict__ adj, int64_t n, int block,
                                          uint32_t* __restrict__ out) {
    const int64_t total = n * block;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t u = i / block;
        const int j = (int)(i - u * block);
        const int v = adj[i];
        uint32_t p01 = 0u, p23 = 0u;
        if (v >= 0) {
            p01 = (bloom_hash(0, (uint32_t)v) & 0xffffu) | (bloom_hash(1, (uint32_t)v) << 16);
            p23 = (bloom_hash(2, (uint32_t)v) & 0xffffu) | (bloom_hash(3, (uint32_t)v) << 16);
        }
        out[u * 2 * block + j] = p01;
        out[u * 2 * block + block + j] = p23;
    }
}

hipError_t launch_visit_table(const int32_t* adj, int64_t n, int block, uint32_t* out, hipStream_t st) {
    if (n < 1 || block < 1) return hipSuccess;
    const int64_t total = n * block;
    const int threads = 256;
    const int64_t blocks = std::min<int64_t>((total + threads - 1) / threads, 65536);
    hipLaunchKernelGGL(visit_table_kernel, dim3((unsigned)blocks), dim3(threads), 0, st, adj, n, block, out);
    return hipGetLastError();
}

}  // namespace song
