// song_kernels.h -- internal interface between the host runtime and the
// search kernels (not part of the public C ABI, which is include/song.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "song_common.h"

namespace song {

// Everything one launch of the search kernel reads or writes.
struct SearchArgs {
    const void* data;        // f32 [n x dstride] (L2/IP) or u64 [n x 4] (Hamming)
    const int32_t* adj;      // [n x block], -1 = empty slot
    const void* queries;     // f32 [nq x dstride] or u64 [nq x 4]
    int32_t* ids_out;        // [nq x k], best first, -1 padded
    float* dists_out;        // [nq x k] or null
    int32_t* trace_out;      // [nq x 12] (QueryTrace field order) or null
    const int32_t* entries;  // [nq x n_entry] entry ids, or null for the evenly spaced seeds
    int64_t n;
    int32_t dstride;         // floats per stored row (dim rounded up to 4)
    int32_t block;           // adjacency slots per node (<= 64)
    int32_t k, ef, nq;
    int32_t bloom_words;     // 64-bit words, power of two
    int32_t bloom_shift;     // log2(bloom_words)
    int32_t n_entry;         // seeds per query (<= 64)
    const int32_t* ks;       // per-query k, or null (uniform k)
    const int32_t* efs;      // per-query queue size, or null (uniform ef); ef := max(ef, k)
    int32_t k_stride;        // output row length (>= every k)
    int32_t* queue_head;     // persistent mode: device counter (zeroed at launch), else null
};

struct LaunchShape {
    int waves_per_block;     // 1, 2 or 4 queries per workgroup
    int persistent_blocks;   // grid size in persistent mode
};

// metric: METRIC_*; S: set slots per lane (cap <= 64*S, S in {1,2,4,8});
// G: lanes per row (power of two); qch: 16-byte chunks per lane (G*qch*4 >= dstride).
hipError_t launch_search(const SearchArgs& a, int metric, int S, int G, int qch, const LaunchShape& shape,
                         hipStream_t st);

}  // namespace song
