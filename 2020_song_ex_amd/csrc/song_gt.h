// song_gt.h -- internal interface of the brute-force ground-truth kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "song_common.h"

namespace song {

struct GtArgs {
    const void* data;     // index rows (f32 [n x dstride] or u64 [n x 4])
    const void* queries;  // same layout, [nq]
    uint64_t* partial;    // [nq x parts x k] keys (workspace)
    int32_t* ids_out;     // [nq x k] device
    float* dists_out;     // [nq x k] device or null
    int64_t n;
    int64_t slice;        // rows per part
    int32_t dstride;
    int32_t nq, k, parts;
};

// qw: queries per wave (1, 2, 4, 8); k <= 128.
hipError_t launch_bruteforce(const GtArgs& a, int metric, int qw, hipStream_t st);

}  // namespace song
