// nsw_build.cpp -- the CPU NSW proximity-graph build.
//
// Algorithm: the reference's sequential insertion, build_kernel_graph
// (kernalGraph.h:140-149, identical to graph.h:190-199) and
// build_hashed_graph (hashKernelGraph.h:196-207):
//   for u = 1..n-1:
//     beam search from node 0 over already-inserted nodes with a frontier
//     min-heap and an EF_CONSTRUCTION-bounded result max-heap
//     (kg_search_layer, kernalGraph.h:77-110);
//     keep the SEARCH_DEGREE nearest (kg_select_neighbors, :112-120);
//     write them as u's row (add_vertex, :32-35);
//     add the reverse edge v->u to each, evicting v's farthest neighbour if
//     the row is full and u is strictly closer (kg_add_reverse_edge, :123-137).
// Heaps are std::priority_queue with the reference's comparators, so ties
// resolve as they do there (the Hamming frontier compares distances only,
// hashKernelGraph.h:135).  The visited set is exact, as unordered_set is.
//
// What is new:
//   * every knob is a parameter (row width, degrees, beam width);
//   * an optional batched schedule: nodes [s, e) are searched in parallel
//     against the graph frozen at s, then their rows and reverse edges are
//     applied in node order (reverse edges grouped by target row, which keeps
//     the per-row order).  batch_div = 0 is the reference's exact sequential
//     build; batch_div = D uses batches of max(1, s / D) after the first
//     seq_prefix nodes.  The result is deterministic for any thread count.
//   * distance order: the sequential float sum of distance.h:5-12 (exact
//     reference order), or 8 independent partial sums (fast); both are exact
//     for integer-valued vectors whose squared distances stay below 2^24.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <queue>
#include <thread>
#include <utility>
#include <vector>

#include "nsw_build.h"

namespace song {

namespace {

// ---------------------------------------------------------------- distances

inline float l2_seq(const float* a, const float* b, int dim) {
    float s = 0.f;
    for (int i = 0; i < dim; i++) {
        float d = a[i] - b[i];
        float p = d * d;
        s = s + p;
    }
    return s;
}

inline float l2_fast(const float* a, const float* b, int dim) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int i = 0;
    for (; i + 8 <= dim; i += 8) {
        for (int j = 0; j < 8; j++) {
            float d = a[i + j] - b[i + j];
            acc[j] += d * d;
        }
    }
    for (; i < dim; i++) {
        float d = a[i] - b[i];
        acc[i & 7] += d * d;
    }
    return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

inline float ip_seq(const float* a, const float* b, int dim) {
    float s = 0.f;
    for (int i = 0; i < dim; i++) {
        float p = a[i] * b[i];
        s = s + p;
    }
    return -s;
}

inline float ip_fast(const float* a, const float* b, int dim) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int i = 0;
    for (; i + 8 <= dim; i += 8)
        for (int j = 0; j < 8; j++) acc[j] += a[i + j] * b[i + j];
    for (; i < dim; i++) acc[i & 7] += a[i] * b[i];
    return -(((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7])));
}

inline int hamming256(const uint64_t* a, const uint64_t* b) {
    return __builtin_popcountll(a[0] ^ b[0]) + __builtin_popcountll(a[1] ^ b[1]) +
           __builtin_popcountll(a[2] ^ b[2]) + __builtin_popcountll(a[3] ^ b[3]);
}

// A distance space: float metrics use pair<float,int> heaps with the pair
// order (kernalGraph.h:15,81-82); Hamming uses pair<int,int> with a
// distance-only frontier comparator (hashKernelGraph.h:17,135-137).
struct FloatSpace {
    using D = float;
    using KP = std::pair<float, int>;
    struct FrontierCmp {
        bool operator()(const KP& a, const KP& b) const { return a > b; }  // std::greater<KP>
    };
    const float* x;
    int dim;
    int metric;
    bool seq;
    D dist(int64_t a, int64_t b) const {
        const float* pa = x + a * (int64_t)dim;
        const float* pb = x + b * (int64_t)dim;
        if (metric == METRIC_IP) return seq ? ip_seq(pa, pb, dim) : ip_fast(pa, pb, dim);
        return seq ? l2_seq(pa, pb, dim) : l2_fast(pa, pb, dim);
    }
    void prefetch(int64_t a) const {
        const char* p = (const char*)(x + a * (int64_t)dim);
        for (int off = 0; off < dim * 4; off += 64) __builtin_prefetch(p + off);
    }
};

struct HammingSpace {
    using D = int;
    using KP = std::pair<int, int>;
    struct FrontierCmp {
        bool operator()(const KP& a, const KP& b) const { return a.first > b.first; }
    };
    const uint64_t* x;
    D dist(int64_t a, int64_t b) const { return hamming256(x + a * 4, x + b * 4); }
    void prefetch(int64_t a) const { __builtin_prefetch(x + a * 4); }
};

template <class Space>
struct Builder {
    using D = typename Space::D;
    using KP = typename Space::KP;
    const Space& sp;
    int64_t n;
    int block, fixed_deg, search_deg, ef_c;
    int32_t* adj;

    // kg_search_layer (kernalGraph.h:77-110) for node u, seeing nodes < lim.
    // Appends the search_deg nearest (kg_select_neighbors, :112-120) to out.
    void search_layer(int u, int lim, std::vector<uint32_t>& stamp, uint32_t& epoch,
                      std::vector<int>& out) const {
        if (++epoch == 0) {  // wrapped: clear stamps
            std::fill(stamp.begin(), stamp.end(), 0u);
            epoch = 1;
        }
        std::priority_queue<KP, std::vector<KP>, typename Space::FrontierCmp> cands;
        std::priority_queue<KP> results;
        const int ep = 0;
        const D ep_d = sp.dist(u, ep);
        cands.push({ep_d, ep});
        results.push({ep_d, ep});
        stamp[ep] = epoch;
        while (!cands.empty()) {
            KP top = cands.top();
            cands.pop();
            if ((int)results.size() >= ef_c && top.first > results.top().first) break;
            const int32_t* row = adj + (int64_t)top.second * block;
            for (int j = 0; j < block; j++) {
                int v = row[j];
                if (v >= 0 && v < lim && stamp[v] != epoch) sp.prefetch(v);
            }
            for (int j = 0; j < block; j++) {
                int v = row[j];
                if (v < 0 || v >= lim || stamp[v] == epoch) continue;
                stamp[v] = epoch;
                D d = sp.dist(u, v);
                cands.push({d, v});
                results.push({d, v});
                if ((int)results.size() > ef_c) results.pop();
            }
        }
        std::vector<KP> ret;
        ret.reserve(results.size());
        while (!results.empty()) {
            ret.push_back(results.top());
            results.pop();
        }
        std::reverse(ret.begin(), ret.end());
        out.clear();
        for (auto& kp : ret) {
            if ((int)out.size() >= search_deg) break;
            out.push_back(kp.second);
        }
    }

    // add_vertex (kernalGraph.h:32-35).
    void add_vertex(int u, const std::vector<int>& nb) {
        int m = std::min((int)nb.size(), fixed_deg);
        for (int j = 0; j < m; j++) adj[(int64_t)u * block + j] = nb[j];
    }

    // kg_add_reverse_edge (kernalGraph.h:123-137).
    void add_reverse_edge(int v, int u) {
        int32_t* row = adj + (int64_t)v * block;
        for (int j = 0; j < fixed_deg; j++) {
            if (row[j] < 0) {
                row[j] = u;
                return;
            }
        }
        D du = sp.dist(v, u);
        int fj = 0;
        D fd = sp.dist(v, row[0]);
        for (int j = 1; j < fixed_deg; j++) {
            D d = sp.dist(v, row[j]);
            if (d > fd) {
                fd = d;
                fj = j;
            }
        }
        if (du < fd) row[fj] = u;
    }
};

void parallel_for(int64_t count, int threads, const std::function<void(int, int64_t)>& fn) {
    if (threads <= 1 || count < 2 * threads) {
        for (int64_t i = 0; i < count; i++) fn(0, i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> th;
    th.reserve(threads);
    for (int t = 0; t < threads; t++) {
        th.emplace_back([&, t]() {
            for (;;) {
                int64_t i = next.fetch_add(1);
                if (i >= count) break;
                fn(t, i);
            }
        });
    }
    for (auto& x : th) x.join();
}

template <class Space>
void run_build(const Space& sp, const BuildParams& p, int32_t* adj) {
    const int64_t n = p.n;
    std::fill(adj, adj + n * p.block, -1);
    Builder<Space> b{sp, n, p.block, p.fixed_deg, p.search_deg, p.ef_construction, adj};
    const int threads = p.n_threads < 1 ? 1 : p.n_threads;
    std::vector<std::vector<uint32_t>> stamps(threads, std::vector<uint32_t>((size_t)n, 0u));
    std::vector<uint32_t> epochs(threads, 0u);
    std::vector<std::vector<int>> nbrs;
    std::vector<std::pair<int, int>> rev;
    auto t0 = std::chrono::steady_clock::now();
    int64_t next_report = n / 10;
    int64_t s = 1;
    while (s < n) {
        int64_t bs = 1;
        if (p.batch_div > 0 && s >= p.seq_prefix) bs = std::max<int64_t>(1, s / p.batch_div);
        const int64_t e = std::min(n, s + bs);
        const int64_t cnt = e - s;
        if ((int64_t)nbrs.size() < cnt) nbrs.resize(cnt);
        const int lim = (int)s;
        if (cnt == 1) {
            b.search_layer((int)s, lim, stamps[0], epochs[0], nbrs[0]);
        } else {
            parallel_for(cnt, threads, [&](int t, int64_t i) {
                b.search_layer((int)(s + i), lim, stamps[t], epochs[t], nbrs[i]);
            });
        }
        for (int64_t i = 0; i < cnt; i++) b.add_vertex((int)(s + i), nbrs[i]);
        if (cnt == 1) {
            for (int v : nbrs[0]) b.add_reverse_edge(v, (int)s);
        } else {
            // reverse edges grouped by target row, each row's edges in node order
            rev.clear();
            for (int64_t i = 0; i < cnt; i++)
                for (int v : nbrs[i]) rev.push_back({v, (int)(s + i)});
            std::stable_sort(rev.begin(), rev.end(),
                             [](const std::pair<int, int>& a, const std::pair<int, int>& c) { return a.first < c.first; });
            std::vector<int64_t> starts;
            for (size_t i = 0; i < rev.size(); i++)
                if (i == 0 || rev[i].first != rev[i - 1].first) starts.push_back((int64_t)i);
            starts.push_back((int64_t)rev.size());
            const int64_t groups = (int64_t)starts.size() - 1;
            parallel_for(groups, threads, [&](int, int64_t g) {
                for (int64_t i = starts[g]; i < starts[g + 1]; i++) b.add_reverse_edge(rev[i].first, rev[i].second);
            });
        }
        s = e;
        if (p.verbose && s >= next_report) {
            double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::fprintf(stderr, "[nsw_build] %lld/%lld nodes inserted, %.1f s\n", (long long)s, (long long)n, sec);
            std::fflush(stderr);
            next_report += n / 10;
        }
    }
}

}  // namespace

int nsw_build(const void* vecs, const BuildParams& p, int32_t* adj) {
    if (p.metric == METRIC_HAMMING) {
        HammingSpace sp{(const uint64_t*)vecs};
        run_build(sp, p, adj);
    } else {
        FloatSpace sp{(const float*)vecs, p.dim, p.metric, p.exact_order != 0};
        run_build(sp, p, adj);
    }
    return 0;
}

}  // namespace song
