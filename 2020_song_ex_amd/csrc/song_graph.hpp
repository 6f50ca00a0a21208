// song_graph.hpp -- the reference's C++ search interface, over libsong's C ABI.
//
// A C++ caller of the reference includes kernalGraph.h / hashKernelGraph.h and
// uses KernelGraph, build_kernel_graph, upload_to_gpu and gpu_search.  This
// header offers the same names with the same argument meaning, so such a
// caller switches by changing its include; the work runs in libsong.so.
// Differences, all deliberate:
//   * errors throw song::Error (the reference ignores CUDA status codes);
//   * ef above 512 throws instead of being clamped to 50 (kernalGraph.h:316);
//   * knobs that are constexpr in config.h are fields of BuildConfig /
//     SearchConfig with the reference values as defaults.
#pragma once
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/song.h"

namespace song {

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void check(int rc, const char* what) {
    if (rc != SONG_OK) throw Error(std::string(what) + ": " + song_last_error());
}

// config.h:9-58
constexpr int BLOCK_SIZE = 32;
constexpr int FIXED_DEGREE = 31;
constexpr int SEARCH_DEGREE = 15;
constexpr int EF_CONSTRUCTION = 150;
constexpr int pq_size = 50;
constexpr int TOPK = 10;
constexpr int N_MULTIPROBE = 4;
constexpr int N_MULTIQUERY = 512;
constexpr int hash_bits = 256;
constexpr int HASH_WORDS = hash_bits / 32;
constexpr int BF_WORDS = 256;

struct BuildConfig {
    int block = BLOCK_SIZE;
    int fixed_degree = FIXED_DEGREE;
    int search_degree = SEARCH_DEGREE;
    int ef_construction = EF_CONSTRUCTION;
    int n_threads = 1;
    int batch_div = 0;          // 0: the reference's sequential insertion
    int64_t seq_prefix = 1024;
    bool exact_order = true;    // distance.h summation order
    bool verbose = false;
};

struct SearchConfig {
    int bloom_bits = BF_WORDS * 64;
    int n_seeds = N_MULTIPROBE;
    int waves_per_block = 0;
    int lanes_per_row = 0;
};

// telemetry.h:5-41
using QueryTrace = song_query_trace;
struct BatchTiming {
    int batch_index = 0, batch_offset = 0, batch_size = 0;
    double h2d_ms = 0, memset_ms = 0, kernel_ms = 0, d2h_ms = 0;
};
struct GpuSearchTrace {
    int ef_search = 0;
    int batch_limit = 0;
    double alloc_ms = 0, h2d_ms = 0, memset_ms = 0, kernel_ms = 0, d2h_ms = 0, total_ms = 0;
    std::vector<BatchTiming> batches;
    std::vector<QueryTrace> queries;
};

namespace detail {
class Resident {
  public:
    Resident() = default;
    Resident(const Resident&) = delete;
    Resident& operator=(const Resident&) = delete;
    Resident(Resident&& o) noexcept : ix_(o.ix_) { o.ix_ = nullptr; }
    Resident& operator=(Resident&& o) noexcept {
        if (this != &o) {
            reset();
            ix_ = o.ix_;
            o.ix_ = nullptr;
        }
        return *this;
    }
    ~Resident() { reset(); }
    void reset() {
        if (ix_) song_index_destroy(ix_);
        ix_ = nullptr;
    }
    song_index* get() const { return ix_; }
    song_index** out() { return &ix_; }

  private:
    song_index* ix_ = nullptr;
};

inline std::vector<int> run_search(song_index* ix, const void* queries, int nq, int k, int batch, int ef,
                                   const SearchConfig& sc, GpuSearchTrace* trace) {
    if (!ix) throw Error("graph not uploaded: call upload_to_gpu() first");
    if (ef < k) ef = k;  // kernalGraph.h:315
    song_search_opts o;
    std::memset(&o, 0, sizeof o);
    o.bloom_bits = sc.bloom_bits;
    o.batch = batch;
    o.n_seeds = sc.n_seeds;
    o.waves_per_block = sc.waves_per_block;
    o.lanes_per_row = sc.lanes_per_row;
    std::vector<int> ids((size_t)nq * k);
    song_timing tm;
    const int nb = batch > 0 ? (nq + batch - 1) / batch : 1;
    std::vector<song_batch_timing> bt(nb > 0 ? nb : 1);
    std::vector<QueryTrace> qt(trace ? nq : 0);
    check(song_search(ix, queries, nq, k, ef, &o, ids.data(), nullptr, trace ? qt.data() : nullptr, &tm, bt.data()),
          "song_search");
    if (trace) {
        *trace = GpuSearchTrace{};
        trace->ef_search = tm.ef_used;
        trace->batch_limit = batch;
        trace->alloc_ms = tm.alloc_ms;
        trace->h2d_ms = tm.h2d_ms;
        trace->memset_ms = tm.memset_ms;
        trace->kernel_ms = tm.kernel_ms;
        trace->d2h_ms = tm.d2h_ms;
        trace->total_ms = tm.total_ms;
        for (int i = 0; i < tm.n_batches; i++) {
            BatchTiming b;
            b.batch_index = bt[i].batch_index;
            b.batch_offset = bt[i].batch_offset;
            b.batch_size = bt[i].batch_size;
            b.h2d_ms = bt[i].h2d_ms;
            b.memset_ms = bt[i].memset_ms;
            b.kernel_ms = bt[i].kernel_ms;
            b.d2h_ms = bt[i].d2h_ms;
            trace->batches.push_back(b);
        }
        trace->queries = std::move(qt);
    }
    return ids;
}
}  // namespace detail

// KernelGraph (kernalGraph.h:20-71): host data + adjacency, device copy after
// upload_to_gpu().
struct KernelGraph {
    int num_nodes = 0;
    int dim = 0;
    int block = BLOCK_SIZE;
    std::vector<float> data;
    std::vector<int> adj;
    detail::Resident gpu;

    KernelGraph(int n, int d, std::vector<float> data_, int block_ = BLOCK_SIZE)
        : num_nodes(n), dim(d), block(block_), data(std::move(data_)), adj((size_t)n * block_, -1) {}

    void upload_to_gpu(int device = 0) {
        gpu.reset();
        check(song_index_create(data.data(), adj.data(), num_nodes, dim, block, SONG_METRIC_L2, device, gpu.out()),
              "song_index_create");
    }
};

inline KernelGraph build_kernel_graph(std::vector<float> data, int n, int dim, const BuildConfig& bc = {}) {
    KernelGraph g(n, dim, std::move(data), bc.block);
    check(song_nsw_build(g.data.data(), n, dim, SONG_METRIC_L2, bc.block, bc.fixed_degree, bc.search_degree,
                         bc.ef_construction, bc.n_threads, bc.batch_div, bc.seq_prefix, bc.exact_order ? 1 : 0,
                         bc.verbose ? 1 : 0, g.adj.data()),
          "song_nsw_build");
    return g;
}

// Graph (graph.h:11-115): the CPU graph and its search, the reference's CPU
// baseline.  search() returns the result ids worst first, as
// Graph::search's vector holds them.
struct Graph {
    int num_nodes = 0;
    int dim = 0;
    int block = BLOCK_SIZE;
    std::vector<float> data;
    std::vector<int> adj;

    Graph(int n, int d, std::vector<float> data_, int block_ = BLOCK_SIZE)
        : num_nodes(n), dim(d), block(block_), data(std::move(data_)), adj((size_t)n * block_, -1) {}

    std::vector<int> search(const float* query, int k, int ef_search = 0, QueryTrace* trace = nullptr) const {
        std::vector<int> ids(k);
        QueryTrace t;
        check(song_cpu_search(data.data(), adj.data(), num_nodes, dim, block, query, 1, k, ef_search, N_MULTIPROBE,
                              1, ids.data(), nullptr, &t, nullptr, nullptr),
              "song_cpu_search");
        ids.resize(t.result_count);
        if (trace) *trace = t;
        return ids;
    }
};

// build_graph (graph.h:190-199): the same NSW build as build_kernel_graph.
inline Graph build_graph(std::vector<float> data, int n, int dim, const BuildConfig& bc = {}) {
    Graph g(n, dim, std::move(data), bc.block);
    check(song_nsw_build(g.data.data(), n, dim, SONG_METRIC_L2, bc.block, bc.fixed_degree, bc.search_degree,
                         bc.ef_construction, bc.n_threads, bc.batch_div, bc.seq_prefix, bc.exact_order ? 1 : 0,
                         bc.verbose ? 1 : 0, g.adj.data()),
          "song_nsw_build");
    return g;
}

// gpu_search (kernalGraph.h:307-382)
inline std::vector<int> gpu_search(const KernelGraph& g, const float* queries, int nq, int k,
                                   int n_multiquery = N_MULTIQUERY, int ef_search = pq_size,
                                   GpuSearchTrace* trace = nullptr, const SearchConfig& sc = {}) {
    return detail::run_search(g.gpu.get(), queries, nq, k, n_multiquery, ef_search, sc, trace);
}

// BitHash (bitHash.h:17-56)
struct BitHash {
    int dim = 0;
    std::vector<float> R;
    explicit BitHash(int d, unsigned seed = 42) : dim(d), R((size_t)hash_bits * d) {
        check(song_bithash_matrix(d, seed, R.data()), "song_bithash_matrix");
    }
    // Writes HASH_WORDS uint32 words (the reference's layout).
    void project(const float* vec, uint32_t* out) const {
        uint64_t c[4];
        check(song_bithash_project(R.data(), dim, vec, 1, 1, c), "song_bithash_project");
        std::memcpy(out, c, sizeof c);
    }
    std::vector<uint32_t> project_all(const float* data, int n, int n_threads = 1) const {
        std::vector<uint32_t> out((size_t)n * HASH_WORDS);
        check(song_bithash_project(R.data(), dim, data, n, n_threads, (uint64_t*)out.data()), "song_bithash_project");
        return out;
    }
    // project_all on GPU `device` (song_bithash_project_device): the same codes bit for bit,
    // host buffers in and out.
    std::vector<uint32_t> project_all_gpu(const float* data, int n, int device = 0) const {
        std::vector<uint32_t> out((size_t)n * HASH_WORDS);
        void *d_R = nullptr, *d_x = nullptr, *d_c = nullptr;
        auto release = [&] {
            song_dev_free(d_R);
            song_dev_free(d_x);
            song_dev_free(d_c);
        };
        int rc = song_dev_alloc(device, (int64_t)R.size() * 4, &d_R);
        if (rc == SONG_OK) rc = song_dev_alloc(device, (int64_t)n * dim * 4, &d_x);
        if (rc == SONG_OK) rc = song_dev_alloc(device, (int64_t)n * 32, &d_c);
        if (rc == SONG_OK) rc = song_memcpy_h2d(d_R, R.data(), (int64_t)R.size() * 4);
        if (rc == SONG_OK) rc = song_memcpy_h2d(d_x, data, (int64_t)n * dim * 4);
        if (rc == SONG_OK)
            rc = song_bithash_project_device((const float*)d_R, dim, (const float*)d_x, n, device, (uint64_t*)d_c,
                                             nullptr);
        if (rc == SONG_OK) rc = song_memcpy_d2h(out.data(), d_c, (int64_t)n * 32);
        const std::string err = rc == SONG_OK ? "" : song_last_error();
        release();
        if (rc != SONG_OK) throw Error(std::string("song_bithash_project_device: ") + err);
        return out;
    }
};

// HashKernelGraph (hashKernelGraph.h:22-126)
struct HashKernelGraph {
    int num_nodes = 0;
    int block = BLOCK_SIZE;
    BitHash bh;
    std::vector<int> adj;
    std::vector<uint32_t> hashed;
    detail::Resident gpu;

    HashKernelGraph(int n, BitHash bh_, std::vector<uint32_t> hashed_, int block_ = BLOCK_SIZE)
        : num_nodes(n), block(block_), bh(std::move(bh_)), adj((size_t)n * block_, -1), hashed(std::move(hashed_)) {}

    void upload_to_gpu(int device = 0) {
        gpu.reset();
        check(song_index_create(hashed.data(), adj.data(), num_nodes, hash_bits, block, SONG_METRIC_HAMMING, device,
                                gpu.out()),
              "song_index_create");
    }
};

inline HashKernelGraph build_hashed_graph(const float* raw_data, int n, int orig_dim, const BuildConfig& bc = {},
                                          unsigned seed = 42) {
    BitHash bh(orig_dim, seed);
    std::vector<uint32_t> hashed = bh.project_all(raw_data, n, bc.n_threads);
    HashKernelGraph g(n, std::move(bh), std::move(hashed), bc.block);
    check(song_nsw_build(g.hashed.data(), n, hash_bits, SONG_METRIC_HAMMING, bc.block, bc.fixed_degree,
                         bc.search_degree, bc.ef_construction, bc.n_threads, bc.batch_div, bc.seq_prefix, 1,
                         bc.verbose ? 1 : 0, g.adj.data()),
          "song_nsw_build");
    return g;
}

// gpu_search (hashKernelGraph.h:360-435); queries are HASH_WORDS uint32 each.
inline std::vector<int> gpu_search(const HashKernelGraph& g, const uint32_t* queries, int nq, int k,
                                   int n_multiquery = N_MULTIQUERY, int ef_search = pq_size,
                                   GpuSearchTrace* trace = nullptr, const SearchConfig& sc = {}) {
    return detail::run_search(g.gpu.get(), queries, nq, k, n_multiquery, ef_search, sc, trace);
}

}  // namespace song
