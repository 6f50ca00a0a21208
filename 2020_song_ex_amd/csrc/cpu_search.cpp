// cpu_search.cpp -- song_cpu_search: the reference's CPU search baseline,
// Graph::search (graph.h:33-115) as bench_cpu times it (bench_cpu.cpp:444-452),
// with the same ids and QueryTrace counters.  It is the CPU method of the
// bench_cpu driver and run_compare's repro_cpu_raw; the GPU search never falls
// back to it.
//
// Semantics kept from Graph::search:
//  * n_seeds evenly spaced entry points p * n / n_seeds (size_t arithmetic),
//    a repeated one skipped;
//  * an exact visited set, an unbounded frontier ordered by (distance, id)
//    ascending, and a result set of the ef best ordered by (distance, id);
//  * the search ends when the frontier is empty, or when the result set is
//    full and the popped distance is strictly worse than its worst distance;
//  * a fresh neighbour always enters the frontier, and the result set when it
//    is not full or the distance is strictly below its worst distance;
//  * squared L2 summed in index order with each square rounded on its own
//    (distance.h:5-12).
// Both orders are total on (distance, id) pairs -- every id enters at most
// once -- so any correct heap pops the same sequence as std::priority_queue.
// Here the pair is one 64-bit key, distance bits above the id: for the
// non-negative distances of squared L2 (NaN is refused at the boundary) the
// integer order of the bits is the float order.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/song.h"

namespace song {
void set_last_error(const std::string& msg);
}

namespace {

using Key = uint64_t;

inline Key make_key(float d, int32_t id) {
    uint32_t b;
    std::memcpy(&b, &d, 4);
    return ((Key)b << 32) | (uint32_t)id;
}
inline float key_dist(Key k) {
    const uint32_t b = (uint32_t)(k >> 32);
    float d;
    std::memcpy(&d, &b, 4);
    return d;
}
inline int32_t key_id(Key k) { return (int32_t)(uint32_t)k; }

// Binary heap over keys; MAX = true keeps the largest on top.
template <bool MAX>
struct KeyHeap {
    std::vector<Key> h;
    static bool above(Key a, Key b) { return MAX ? a > b : a < b; }
    void clear() { h.clear(); }
    size_t size() const { return h.size(); }
    bool empty() const { return h.empty(); }
    Key top() const { return h[0]; }
    void push(Key k) {
        size_t i = h.size();
        h.push_back(k);
        while (i > 0) {
            const size_t parent = (i - 1) / 2;
            if (!above(k, h[parent])) break;
            h[i] = h[parent];
            i = parent;
        }
        h[i] = k;
    }
    Key pop() {
        const Key out = h[0];
        const Key last = h.back();
        h.pop_back();
        const size_t n = h.size();
        if (n > 0) {
            size_t i = 0;
            for (;;) {
                size_t c = 2 * i + 1;
                if (c >= n) break;
                if (c + 1 < n && above(h[c + 1], h[c])) c++;
                if (!above(h[c], last)) break;
                h[i] = h[c];
                i = c;
            }
            h[i] = last;
        }
        return out;
    }
};

// Squared L2 in index order, each square rounded before it is added
// (distance.h:5-12), for up to four rows at once: the rows' sums are
// independent chains, so interleaving them keeps each one's order.
inline void l2_rows(const float* q, const float* const* rows, int nr, int dim, float* out) {
    if (nr == 4) {
        float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
        const float *a = rows[0], *b = rows[1], *c = rows[2], *d = rows[3];
        for (int i = 0; i < dim; i++) {
            const float e0 = q[i] - a[i], e1 = q[i] - b[i], e2 = q[i] - c[i], e3 = q[i] - d[i];
            const float p0 = e0 * e0, p1 = e1 * e1, p2 = e2 * e2, p3 = e3 * e3;
            s0 += p0;
            s1 += p1;
            s2 += p2;
            s3 += p3;
        }
        out[0] = s0;
        out[1] = s1;
        out[2] = s2;
        out[3] = s3;
        return;
    }
    for (int r = 0; r < nr; r++) {
        float s = 0.f;
        const float* x = rows[r];
        for (int i = 0; i < dim; i++) {
            const float e = q[i] - x[i];
            const float p = e * e;
            s += p;
        }
        out[r] = s;
    }
}

struct Searcher {
    const float* vecs;
    const int32_t* adj;
    int64_t n;
    int dim, block, n_seeds;
    std::vector<uint32_t> stamp;  // visited: stamp[id] == epoch
    uint32_t epoch = 0;
    KeyHeap<false> frontier;
    KeyHeap<true> results;
    std::vector<int32_t> fresh;
    std::vector<float> fresh_d;

    Searcher(const float* v, const int32_t* a, int64_t n_, int dim_, int block_, int seeds)
        : vecs(v), adj(a), n(n_), dim(dim_), block(block_), n_seeds(seeds), stamp((size_t)n_, 0u) {
        fresh.reserve(block_);
        fresh_d.resize(block_ + 4);
    }

    bool visit(int32_t id) {
        if (stamp[id] == epoch) return false;
        stamp[id] = epoch;
        return true;
    }

    void distances(const float* q, int cnt) {
        const float* rows[4];
        int i = 0;
        for (; i + 4 <= cnt; i += 4) {
            for (int r = 0; r < 4; r++) rows[r] = vecs + (size_t)fresh[i + r] * dim;
            l2_rows(q, rows, 4, dim, fresh_d.data() + i);
        }
        const int rest = cnt - i;
        for (int r = 0; r < rest; r++) rows[r] = vecs + (size_t)fresh[i + r] * dim;
        if (rest) l2_rows(q, rows, rest, dim, fresh_d.data() + i);
    }

    // Graph::search of one query; ids worst first (the order its result
    // vector has), padded with -1.
    void search(const float* q, int k, int ef, int32_t* ids, float* dists, song_query_trace* tr) {
        if (ef < k) ef = k;
        if (++epoch == 0) {  // wrapped: forget every stamp
            std::fill(stamp.begin(), stamp.end(), 0u);
            epoch = 1;
        }
        frontier.clear();
        results.clear();
        song_query_trace t;
        std::memset(&t, 0, sizeof t);
        for (int p = 0; p < n_seeds; p++) {
            const int32_t ep = (int32_t)((size_t)p * (size_t)n / (size_t)n_seeds);
            if (!visit(ep)) continue;
            fresh.assign(1, ep);
            distances(q, 1);
            const Key key = make_key(fresh_d[0], ep);
            frontier.push(key);
            results.push(key);
            t.seed_count++;
            t.visited_nodes++;
            t.distance_evals++;
            t.frontier_pushes++;
            t.result_updates++;
            t.peak_frontier = std::max(t.peak_frontier, (int32_t)frontier.size());
            t.peak_results = std::max(t.peak_results, (int32_t)results.size());
        }
        while (!frontier.empty()) {
            const Key cur = frontier.pop();
            t.expanded_nodes++;
            if ((int)results.size() == ef && key_dist(cur) > key_dist(results.top())) {
                t.early_stopped = 1;
                break;
            }
            const int32_t* row = adj + (size_t)key_id(cur) * block;
            fresh.clear();
            for (int j = 0; j < block; j++) {
                t.neighbor_slots_scanned++;
                const int32_t v = row[j];
                if (v < 0 || !visit(v)) continue;
                fresh.push_back(v);
                t.visited_nodes++;
                t.candidate_nodes++;
            }
            const int cnt = (int)fresh.size();
            distances(q, cnt);
            t.distance_evals += cnt;
            for (int i = 0; i < cnt; i++) {
                const float d = fresh_d[i];
                const Key key = make_key(d, fresh[i]);
                frontier.push(key);
                t.frontier_pushes++;
                t.peak_frontier = std::max(t.peak_frontier, (int32_t)frontier.size());
                if ((int)results.size() < ef || d < key_dist(results.top())) {
                    results.push(key);
                    if ((int)results.size() > ef) results.pop();
                    t.result_updates++;
                    t.peak_results = std::max(t.peak_results, (int32_t)results.size());
                }
            }
        }
        while ((int)results.size() > k) results.pop();
        int cnt = 0;
        while (!results.empty()) {
            const Key key = results.pop();
            ids[cnt] = key_id(key);
            if (dists) dists[cnt] = key_dist(key);
            cnt++;
        }
        for (int i = cnt; i < k; i++) {
            ids[i] = -1;
            if (dists) dists[i] = INFINITY;
        }
        t.result_count = cnt;
        if (tr) *tr = t;
    }
};

}  // namespace

extern "C" int song_cpu_search(const float* vecs, const int32_t* adj, int64_t n, int32_t dim, int32_t block,
                               const float* queries, int32_t nq, int32_t k, int32_t ef, int32_t n_seeds,
                               int32_t n_threads, int32_t* ids_out, float* dists_out, song_query_trace* trace_out,
                               double* per_query_ms_out, double* search_ms_out) {
    auto bad = [](const char* m) {
        song::set_last_error(m);
        return SONG_ERR_ARG;
    };
    if (!vecs || !adj || (nq > 0 && (!queries || !ids_out))) return bad("null buffer");
    if (n < 1 || n > INT32_MAX) return bad("n must be in [1, 2^31)");
    if (dim < 1) return bad("dim must be >= 1");
    if (block < 1 || block > 1024) return bad("block must be in [1, 1024]");
    if (nq < 0) return bad("nq must be >= 0");
    if (k < 1) return bad("k must be >= 1");
    if (n_seeds == 0) n_seeds = 4;  // N_MULTIPROBE, config.h:34
    if (n_seeds < 1 || n_seeds > 64) return bad("n_seeds must be in [1, 64]");
    for (int64_t i = 0; i < n * block; i++)
        if (adj[i] < -1 || adj[i] >= n) return bad("adjacency entry out of range");
    for (int64_t i = 0; i < n * dim; i++)
        if (!std::isfinite(vecs[i])) return bad("vectors must be finite (NaN or Inf found)");
    for (int64_t i = 0; i < (int64_t)nq * dim; i++)
        if (!std::isfinite(queries[i])) return bad("queries must be finite (NaN or Inf found)");
    const int threads = std::max(1, std::min<int>(n_threads, std::max(1, nq)));
    std::atomic<int32_t> next{0};
    auto work = [&]() {
        Searcher s(vecs, adj, n, dim, block, n_seeds);
        for (;;) {
            const int32_t i = next.fetch_add(1);
            if (i >= nq) break;
            const auto t0 = std::chrono::steady_clock::now();
            s.search(queries + (size_t)i * dim, k, ef, ids_out + (size_t)i * k,
                     dists_out ? dists_out + (size_t)i * k : nullptr, trace_out ? trace_out + i : nullptr);
            if (per_query_ms_out)
                per_query_ms_out[i] =
                    std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    };
    const auto t_start = std::chrono::steady_clock::now();
    try {
        if (threads == 1) {
            work();
        } else {
            std::vector<std::thread> pool;
            for (int t = 0; t < threads; t++) pool.emplace_back(work);
            for (auto& th : pool) th.join();
        }
    } catch (const std::bad_alloc&) {
        song::set_last_error("host allocation failed in cpu search");
        return SONG_ERR_ALLOC;
    }
    if (search_ms_out)
        *search_ms_out = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return SONG_OK;
}
