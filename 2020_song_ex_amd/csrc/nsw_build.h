// nsw_build.h -- internal interface of the CPU NSW build.
#pragma once
#include <stdint.h>

#include "song_common.h"

namespace song {

struct BuildParams {
    int64_t n;
    int dim;              // floats per vector (ignored for Hamming: 256-bit codes)
    int metric;
    int block;            // slots per row (BLOCK_SIZE)
    int fixed_deg;        // max edges per row (FIXED_DEGREE, < block)
    int search_deg;       // forward edges of a new node (SEARCH_DEGREE)
    int ef_construction;  // beam width (EF_CONSTRUCTION)
    int n_threads;
    int batch_div;        // 0: exact sequential build; D: batches of s / D
    int64_t seq_prefix;   // nodes inserted one by one before batching starts
    int exact_order;      // 1: distance.h summation order; 0: 8 partial sums
    int verbose;
};

// adj: [n x block] written in full (-1 = empty).  Arguments are validated by
// the caller.
int nsw_build(const void* vecs, const BuildParams& p, int32_t* adj);

}  // namespace song
