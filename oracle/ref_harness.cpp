// ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A C-ABI shim around the reference's own CPU sources, compiled straight from
// where they lie (oracle/Makefile passes -I to the reference checkout; nothing
// from it is copied here).  It exposes:
//   - build_graph        graph.h:190-199   (CPU NSW build, BLOCK_SIZE=32,
//                                          FIXED_DEGREE=31, SEARCH_DEGREE=15,
//                                          EF_CONSTRUCTION=150 from config.h)
//   - Graph::search      graph.h:33-115    (the bench_cpu search, bench_cpu.cpp:444-452)
//   - l2                 distance.h:5-12
//   - hamming, BitHash   bitHash.h:8-56
// These are the checks the oracle and the product's NSW build are pinned
// against, and Graph::search is the timed CPU baseline of bench.py.
// The CUDA half of the reference (kernalGraph.h, hashKernelGraph.h, gpuHeap.h,
// bloomFilter.h) needs cuda_runtime.h and cannot be built here.
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "graph.h"
#include "bitHash.h"

extern "C" {

int ref_block_size(void) { return BLOCK_SIZE; }
int ref_fixed_degree(void) { return FIXED_DEGREE; }
int ref_search_degree(void) { return SEARCH_DEGREE; }
int ref_ef_construction(void) { return EF_CONSTRUCTION; }
int ref_n_multiprobe(void) { return N_MULTIPROBE; }

float ref_l2(const float* a, const float* b, int dim) { return l2(a, b, dim); }

int ref_hamming(const uint32_t* a, const uint32_t* b) { return hamming(a, b); }

// Sequential NSW build of graph.h; adj_out is [n x BLOCK_SIZE].
int ref_build_graph(const float* data, int n, int dim, int32_t* adj_out) {
    std::vector<dist_t> v(data, data + (size_t)n * dim);
    Graph g = build_graph(std::move(v), n, dim);
    std::memcpy(adj_out, g.adj.data(), (size_t)n * BLOCK_SIZE * sizeof(int32_t));
    return 0;
}

// Graph::search over nq queries with n_threads threads (queries split in
// contiguous ranges).  ids_out [nq x k] worst-first as Graph::search returns
// them, -1 padded; trace_out [nq x 12] in telemetry.h order; per_query_ms
// optional.  Returns the wall time in ms of the search loop.
double ref_graph_search(const float* data, const int32_t* adj, int n, int dim, const float* queries,
                        int nq, int k, int ef, int n_threads, int32_t* ids_out, int32_t* trace_out,
                        double* per_query_ms) {
    Graph g(n, dim, std::vector<dist_t>(data, data + (size_t)n * dim));
    std::memcpy(g.adj.data(), adj, (size_t)n * BLOCK_SIZE * sizeof(int32_t));
    if (n_threads < 1) n_threads = 1;
    using clk = std::chrono::high_resolution_clock;
    auto work = [&](int lo, int hi) {
        for (int i = lo; i < hi; i++) {
            QueryTrace tr;
            auto t0 = clk::now();
            std::vector<idx_t> r = g.search(queries + (size_t)i * dim, k, ef, &tr);
            auto t1 = clk::now();
            if (per_query_ms) per_query_ms[i] = std::chrono::duration<double, std::milli>(t1 - t0).count();
            for (int j = 0; j < k; j++) ids_out[(size_t)i * k + j] = j < (int)r.size() ? r[j] : -1;
            if (trace_out) {
                int32_t* t = trace_out + (size_t)i * 12;
                t[0] = tr.seed_count; t[1] = tr.visited_nodes; t[2] = tr.expanded_nodes;
                t[3] = tr.neighbor_slots_scanned; t[4] = tr.candidate_nodes; t[5] = tr.distance_evals;
                t[6] = tr.frontier_pushes; t[7] = tr.result_updates; t[8] = tr.peak_frontier;
                t[9] = tr.peak_results; t[10] = tr.result_count; t[11] = tr.early_stopped;
            }
        }
    };
    auto t0 = clk::now();
    if (n_threads == 1) {
        work(0, nq);
    } else {
        std::vector<std::thread> th;
        int per = (nq + n_threads - 1) / n_threads;
        for (int t = 0; t < n_threads; t++) {
            int lo = t * per, hi = std::min(nq, lo + per);
            if (lo < hi) th.emplace_back(work, lo, hi);
        }
        for (auto& x : th) x.join();
    }
    auto t1 = clk::now();
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

// BitHash(dim, seed) projection of n vectors into 8 x uint32 codes each.
int ref_bithash_project(int dim, unsigned seed, const float* vecs, int n, uint32_t* out) {
    BitHash bh(dim, seed);
    for (int i = 0; i < n; i++) bh.project(vecs + (size_t)i * dim, out + (size_t)i * HASH_WORDS);
    return 0;
}

// The projection matrix itself, [hash_bits x dim].
int ref_bithash_matrix(int dim, unsigned seed, float* out) {
    BitHash bh(dim, seed);
    std::memcpy(out, bh.R.data(), bh.R.size() * sizeof(float));
    return 0;
}

}  // extern "C"
