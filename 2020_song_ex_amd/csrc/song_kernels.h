// song_kernels.h -- internal interface between the host runtime and the
// search kernels (not part of the public C ABI, which is include/song.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "song_common.h"

namespace song {

// Learned entry and termination (2 x 64 hidden FP32 MLPs).
constexpr int LM_P = 16;  // query projection dims (entry MLP input)
constexpr int LM_C = 64;  // candidate entry vertices (entry MLP outputs)
constexpr int LM_F = 8;   // per-hop features (stop MLP input)
constexpr int LM_H = 64;  // hidden width (one lane per unit)
constexpr int LM_STOP_FLOATS = LM_H * LM_F + LM_H + LM_H * LM_H + LM_H + LM_H;  // w1 b1 w2 b2 w3

constexpr int LM_COST_FLOATS = LM_H * LM_P + LM_H + LM_H * LM_H + LM_H + LM_H + 1;  // ... w3 b3

// Stop-network weights (w1 [64 x 8], b1, w2 [64 x 64] row-major, b2, w3) into
// the search kernel's layout: w2 interleaved so that element (j, i) sits at
// w2 + (i / 4) * 256 + j * 4 + i % 4 (lane j reads coalesced float4s).
inline void stop_weights_to_kernel_layout(const float* src, float* dst) {
    const int w2 = LM_H * LM_F + LM_H;
    for (int i = 0; i < LM_STOP_FLOATS; i++) dst[i] = src[i];
    for (int j = 0; j < LM_H; j++)
        for (int i = 0; i < LM_H; i++) dst[w2 + (i / 4) * 256 + j * 4 + (i % 4)] = src[w2 + j * LM_H + i];
}

// The query-side networks (entry, cost), run by the planning kernel before
// the search: z = scale * (proj (q - mu)) feeds both.
struct PlanArgs {
    const void* queries;      // as SearchArgs::queries
    int32_t nq;
    int32_t dstride;
    int64_t n;
    const float* mu;          // [dstride] (float metrics; unused for Hamming)
    const float* proj;        // [LM_P x proj_dim] row-major, proj_dim = dstride or 256 (Hamming, +-1 bits)
    const float* proj_scale;  // [LM_P]
    const float* e_w1;        // [LM_H x LM_P]
    const float* e_b1;        // [LM_H]
    const float* e_w2;        // [LM_H x LM_H]
    const float* e_b2;        // [LM_H]
    const float* e_w3;        // [LM_C x LM_H]
    const float* e_b3;        // [LM_C]
    const int32_t* pivots;    // [LM_C] node ids (< n, checked on the host)
    const float* c_weights;   // LM_COST_FLOATS (16 -> 64 -> 64 -> 1), or null
    int32_t n_entry;          // seeds per query (row length of entries)
    int32_t n_ref;            // leading reference seeds (p * n / n_ref) kept in each row, 0 = none
    int32_t* entries;         // [nq x n_entry] out, or null (no entry network)
    float* cost;              // [nq] out (predicted standardised log2 hops), or null
};

// The stop network, read by the search kernel.
struct LearnedArgs {
    // logit = w3 . relu(W2 relu(W1 f + b1) + b2) + b3, features f (LM_F)
    const float* s_weights;   // LM_STOP_FLOATS: w1 [H x F], b1 [H], w2 [H x H], b2 [H], w3 [H]
    float s_b3;
    float stop_logit;         // stop when logit > stop_logit
    float dist_scale;         // feature normalisation of distance differences
    int32_t use_stop;
    int32_t min_hops;         // first hop at which the stop MLP is consulted
    int32_t interval;         // consult every `interval` hops
    int32_t late_hops;        // from this hop on consult every late_interval hops (0 = never)
    int32_t late_interval;
    int32_t max_hops;         // stop after this many expansions (0 = no cap)
    int32_t confirm;          // consecutive consults above the threshold needed to stop (>= 1)
    int32_t result_cap;       // result-set capacity below ef, without the reference's early stop (0 = ef)
    int32_t entry_late_hops;  // > 0: start at the reference seeds, the entry list joins at this hop
    // training traces (all null outside trace mode)
    float* trace;             // [nq x max_evals x (LM_F + 1)]: features, k-th best distance
    int32_t* trace_count;     // [nq]
    float* final_dk;          // [nq] k-th best distance at the end of the search
    int32_t max_evals;
};

// The request ring of the resident serving kernel.  The host writes request r
// into slot r % ring of host-visible arrays and then publishes tail = r + 1;
// waves claim requests with a compare-and-swap on head (device memory), run
// them, write the results into the slot and set done[slot] = r + 1.  Every
// wait is bounded: a wave leaves once stop is set or after idle_ticks without
// finding a request, so the grid always drains on its own.
struct ResidentQueue {
    int64_t* head;        // device memory: next request number to claim
    const int64_t* tail;  // host memory: requests published so far
    const int32_t* stop;  // host memory: nonzero asks every wave to leave
    int64_t* done;        // host memory [ring]: request number + 1 once written
    int32_t ring;         // slots
    int64_t idle_ticks;   // 100 MHz real-time-clock ticks
};

// Everything one launch of the search kernel reads or writes.
struct SearchArgs {
    const void* data;        // f32 [n x dstride] (L2/IP) or u64 [n x 4] (Hamming)
    const int32_t* adj;      // [n x block], -1 = empty slot
    // Visited-filter positions of every adjacency slot's id, hashed once at
    // index creation (launch_visit_table): [n x 2 x block], plane 0 holds
    // hashes 0 | 1 << 16 and plane 1 hashes 2 | 3 << 16 (each the hash's low
    // 16 bits, masked to the filter size at use).  Null: hash in the search.
    // Read for rows of at most 32 slots.
    const uint32_t* vpos;
    const void* queries;     // f32 [nq x dstride] or u64 [nq x 4]
    int32_t* ids_out;        // [nq x k], best first, -1 padded
    float* dists_out;        // [nq x k] or null
    int32_t* trace_out;      // [nq x 12] (QueryTrace field order) or null
    const int32_t* entries;  // [nq x n_entry] entry ids, or null for the evenly spaced seeds
    int64_t n;
    int32_t dstride;         // floats per stored row (dim rounded up to 4)
    int32_t off32;           // 1: n < 2^23 and n * row bytes < 2^31 (32-bit row offsets)
    int32_t block;           // adjacency slots per node (<= 64)
    int32_t k, ef, nq;
    int32_t bloom_words;     // 64-bit words, power of two
    int32_t bloom_shift;     // log2(bloom_words)
    int32_t n_entry;         // seeds per query (<= 64)
    const int32_t* ks;       // per-query k, or null (uniform k)
    const int32_t* efs;      // per-query queue size, or null (uniform ef); ef := max(ef, k)
    int32_t k_stride;        // output row length (>= every k)
    int32_t* queue_head;     // persistent mode: device counter (zeroed at launch), else null
    const int32_t* order;    // [nq] queue position -> query index, or null (identity)
    const ResidentQueue* resident;  // device copy of the serving ring, or null
    int32_t learned;         // 1: use L (stop / trace as its flags say)
    LearnedArgs L;
};

struct LaunchShape {
    int waves_per_block;     // 1, 2, 4 or 8 queries per workgroup
    int persistent_blocks;   // grid size in persistent mode
};

// Forward passes of the kernels' own inference code on given inputs:
// which = 0: stop network (L.s_weights) on X [n x 8] -> Y [n];
// which = 1: entry network (P.e_*) on X [n x 16] -> Y [n x 64];
// which = 2: cost network (P.c_weights) on X [n x 16] -> Y [n].
hipError_t launch_learned_forward(const LearnedArgs& L, const PlanArgs& P, int which, const float* X, int64_t n,
                                  float* Y, hipStream_t st);

// Planning pass of the learned search: entry seeds and predicted cost per
// query (one wave per query), then, when order is non-null, the queue order
// by predicted cost, most expensive first (counting sort over 256 bins).
hipError_t launch_plan(const PlanArgs& P, int metric, int32_t* order, hipStream_t st);

// metric: METRIC_*; S: set slots per lane (cap <= 64*S, S = 1..8);
// G: lanes per row (power of two); qch: 16-byte chunks per lane (G*qch*4 >= dstride).
hipError_t launch_search(const SearchArgs& a, int metric, int S, int G, int qch, const LaunchShape& shape,
                         hipStream_t st);
// The visited-filter position table of SearchArgs::vpos for an adjacency
// [n x block] already in device memory (out: n * 2 * block words).
hipError_t launch_visit_table(const int32_t* adj, int64_t n, int block, uint32_t* out, hipStream_t st);
// BitHash codes of n vectors (uint64 [n x 4]) from R [256 x dim], dim <= 2048,
// all in device memory: the sequential float dot products of song_bithash_project.
hipError_t launch_bithash(const float* R, int dim, const float* X, int64_t n, uint64_t* codes, hipStream_t st);

}  // namespace song
