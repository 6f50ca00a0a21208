/*
 * song.h -- C ABI of libsong.so, the MI355X-native SONG graph search.
 *
 * The reference exposes its search as a C++ header API (KernelGraph /
 * HashKernelGraph, build_kernel_graph, upload_to_gpu, gpu_search) that its
 * bench_gpu driver calls in-process, and run_compare.py drives bench_gpu as a
 * subprocess.  These entry points replace that header API one for one; each
 * declaration cites the reference interface it replaces.  No C++ types or
 * exceptions cross the boundary: plain pointers, sizes and int status codes.
 *
 * Conventions
 *   - Every call returns 0 (SONG_OK) or a negative SONG_ERR_* code;
 *     song_last_error() then holds a message (thread-local).
 *   - The caller owns every host buffer.  The library owns the device copy
 *     of the index (song_index_create / song_index_destroy).
 *   - One song_index is used by one host thread at a time; its work is
 *     serialised on the index's own HIP stream.
 *   - Out-of-range arguments are errors, never silently clamped.  The one
 *     rule kept from the reference is ef := max(ef, k) (kernalGraph.h:315);
 *     its ef <= 50 cap (kernalGraph.h:316) is lifted to ef <= 512.
 *   - Metrics: L2 = squared Euclidean; IP = inner product stored as -<q,x>
 *     (smaller is better everywhere); HAMMING = popcount distance on 256-bit
 *     codes stored as 4 x uint64 (the bits of the reference's 8 x uint32).
 */
#ifndef SONG_H
#define SONG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SONG_ABI_VERSION 2

enum {
    SONG_METRIC_L2 = 0,
    SONG_METRIC_IP = 1,
    SONG_METRIC_HAMMING = 2
};

enum {
    SONG_OK = 0,
    SONG_ERR_ARG = -1,   /* an argument is out of range or null */
    SONG_ERR_HIP = -2,   /* the HIP runtime reported an error */
    SONG_ERR_ALLOC = -3, /* a device or host allocation failed */
    SONG_ERR_IO = -4,    /* a file could not be read or written */
    SONG_ERR_STATE = -5  /* the call does not fit the object's state */
};

typedef struct song_index song_index;

/* Per-query counters, field for field QueryTrace (telemetry.h:5-18). */
typedef struct song_query_trace {
    int32_t seed_count;
    int32_t visited_nodes;
    int32_t expanded_nodes;
    int32_t neighbor_slots_scanned;
    int32_t candidate_nodes;
    int32_t distance_evals;
    int32_t frontier_pushes;
    int32_t result_updates;
    int32_t peak_frontier;
    int32_t peak_results;
    int32_t result_count;
    int32_t early_stopped;
} song_query_trace;

/* Timing of one search call, the fields of GpuSearchTrace (telemetry.h:30-41)
 * without the per-batch and per-query vectors.  h2d/kernel/d2h are measured
 * with HIP events on the index stream; total_ms is host wall time of the
 * call's batch loop (allocation excluded), as gpu_search's total_ms
 * (kernalGraph.h:341,376). */
typedef struct song_timing {
    double alloc_ms;
    double h2d_ms;
    double memset_ms;
    double kernel_ms;
    double d2h_ms;
    double total_ms;
    int32_t n_batches;
    int32_t ef_used;
} song_timing;

/* Per-batch timing, BatchTiming (telemetry.h:20-28). */
typedef struct song_batch_timing {
    int32_t batch_index;
    int32_t batch_offset;
    int32_t batch_size;
    int32_t pad_;
    double h2d_ms;
    double memset_ms;
    double kernel_ms;
    double d2h_ms;
} song_batch_timing;

/* Search options; zero-initialise and set what you need.  All knobs are
 * compile-time constants in the reference (config.h:9-58). */
typedef struct song_search_opts {
    int32_t bloom_bits;      /* visited-set bits, power of two in [2^9, 2^16];
                                0 = 16384 (BF_WORDS*64, config.h:56) */
    int32_t batch;           /* queries per kernel launch; 0 = all at once
                                (N_MULTIQUERY, config.h:38, is 512) */
    int32_t n_seeds;         /* evenly spaced entry points, 1..64;
                                0 = 4 (N_MULTIPROBE, config.h:34) */
    int32_t waves_per_block; /* 1, 2, 4 or 8 queries per workgroup; 0 = 4
                                (8 with the learned stop, whose network each
                                workgroup stages in LDS) */
    int32_t lanes_per_row;   /* lanes that share one vector row (power of
                                two, 2..64); 0 = chosen from dim.  Fixes the
                                float summation order (xor butterfly over
                                the group, see song_search_impl.h). */
    int32_t mode;            /* SONG_MODE_PLAIN: one wave per query;
                                SONG_MODE_PERSISTENT: a fixed grid drains a
                                device query queue.  Same results. */
    int32_t waves_per_cu;    /* persistent grid: waves per compute unit,
                                1..32; 0 = 32 */
    int32_t learned;         /* bit 0: entry vertices from the index's entry
                                network; bit 1: stop when the stop network's
                                logit exceeds stop_threshold; bit 2: run the
                                queries in order of the cost network's
                                prediction, most expensive first (results are
                                unchanged, only when each query starts);
                                bit 3 (with bit 0): keep the n_seeds reference
                                seeds and add n_seeds learned ones.
                                Needs a model (song_index_set_model; bit 2 a
                                model with a cost network), k <= 64,
                                dim <= 1024. */
    float stop_threshold;    /* logit threshold of the stop network */
    int32_t stop_min_hops;   /* first hop at which it is consulted */
    int32_t stop_interval;   /* consult every this many hops (>= 1) */
    int32_t stop_max_hops;   /* with learned bit 1: also stop once a query has
                                expanded this many nodes (0 = no cap) */
    int32_t stop_confirm;    /* stop after this many consecutive consults above
                                the threshold (0 = 1) */
    int32_t result_cap;      /* with learned bit 1 and a hop cap: result-set
                                capacity, k <= result_cap < ef (0 = ef).  The
                                search then ends by the stop network, the hop
                                cap or an empty frontier, without the
                                reference's "best candidate worse than the
                                worst of the ef results" test, which needs an
                                ef-sized result set (kernalGraph.h:239-241);
                                the k best are kept the same way.  No
                                reference counterpart. */
    int32_t stop_late_hops;  /* with learned bit 1: from this hop on the stop
                                network is consulted every stop_late_interval
                                hops instead (0 = never; no reference
                                counterpart) */
    int32_t stop_late_interval; /* >= 1 when stop_late_hops > 0 */
    int32_t entry_late_hops; /* with learned bits 0 and 1 (not bit 3) and a
                                queue of at most 64 after ef = max(ef, k): the
                                search starts at the n_seeds reference seeds
                                and the entry network's n_seeds seeds join it
                                at this hop, tested and pushed like a row
                                before the popped node's row, if the search
                                is still running (0 = the entry seeds replace
                                the reference ones from the start; no
                                reference counterpart) */
} song_search_opts;

enum { SONG_MODE_PLAIN = 0, SONG_MODE_PERSISTENT = 1 };

/* Version of this header's ABI (SONG_ABI_VERSION). */
int song_abi_version(void);

/* Message of the last error on this thread ("" if none). */
const char* song_last_error(void);

/* Number of visible HIP devices. */
int song_device_count(int32_t* out);

/* ---------------------------------------------------------------- build */

/* CPU NSW build.  Replaces build_kernel_graph (kernalGraph.h:140-149,
 * = build_graph, graph.h:190-199) and build_hashed_graph's graph part
 * (hashKernelGraph.h:196-207); parameters are config.h's BLOCK_SIZE,
 * FIXED_DEGREE, SEARCH_DEGREE and EF_CONSTRUCTION (config.h:9-24).
 *   vecs: f32 [n x dim] (L2, IP) or uint64 [n x 4] (HAMMING, dim = 256)
 *   batch_div: 0 = the reference's exact sequential insertion;
 *              D > 0 = batched insertion, batch size s / D after the first
 *              seq_prefix nodes (deterministic for any n_threads)
 *   exact_order: 1 = distance.h:5-12 summation order, 0 = 8 partial sums
 *   adj_out: [n x block], caller-owned; -1 marks empty slots. */
int song_nsw_build(const void* vecs, int64_t n, int32_t dim, int32_t metric, int32_t block,
                   int32_t fixed_deg, int32_t search_deg, int32_t ef_construction,
                   int32_t n_threads, int32_t batch_div, int64_t seq_prefix, int32_t exact_order,
                   int32_t verbose, int32_t* adj_out);

/* Graph file (no counterpart in the reference, which rebuilds every run,
 * generate_report.py:625-627): header + int32 adjacency. */
int song_graph_save(const char* path, const int32_t* adj, int64_t n, int32_t block, int32_t metric,
                    int32_t dim);
int song_graph_load_header(const char* path, int64_t* n, int32_t* block, int32_t* metric,
                           int32_t* dim);
int song_graph_load(const char* path, int32_t* adj_out, int64_t n, int32_t block);

/* 1-bit Gaussian random projection (BitHash, bitHash.h:17-56): the
 * [256 x dim] matrix from std::mt19937(seed) + normal_distribution<float>
 * (the reference's query/base hashing uses seed 42, bitHash.h:22), and the
 * projection of n vectors to 256-bit codes (BitHash::project, bitHash.h:32-44;
 * project_all, :50-55) as 4 x uint64 per vector. */
int song_bithash_matrix(int32_t dim, uint32_t seed, float* out);
int song_bithash_project(const float* R, int32_t dim, const float* vecs, int64_t n, int32_t n_threads,
                         uint64_t* codes_out);
/* The same projection on the GPU, the same codes bit for bit (each dot product
 * summed in order in float, each product rounded on its own; no fused or
 * reordered sums).  Replaces BitHash::project / project_all (bitHash.h:32-55)
 * on the Hamming query path.  Every buffer is in device memory of `device`:
 * d_R [256 x dim] (song_bithash_matrix's layout), d_vecs [n x dim] f32,
 * d_codes_out [n x 4] uint64.  dim in [1, 2048].  Synchronous; kernel_ms_out
 * (nullable) gets the kernel's time. */
int song_bithash_project_device(const float* d_R, int32_t dim, const float* d_vecs, int64_t n, int32_t device,
                                uint64_t* d_codes_out, double* kernel_ms_out);

/* ---------------------------------------------------------------- index */

/* Device-resident index: vectors (or codes) and adjacency copied to HBM.
 * Replaces the KernelGraph / HashKernelGraph constructors plus
 * upload_to_gpu (kernalGraph.h:29-43, hashKernelGraph.h:31-33,93-98).
 * Every adjacency entry must be -1 or a node id < n (checked).
 * block <= 64.  dim: floats per vector (L2, IP) or 256 (HAMMING). */
int song_index_create(const void* vecs, const int32_t* adj, int64_t n, int32_t dim, int32_t block,
                      int32_t metric, int32_t device, song_index** out);

/* Frees the device copy (~KernelGraph, kernalGraph.h:45). */
void song_index_destroy(song_index* index);

/* The index's visit table: for every adjacency slot, the visited-filter
 * positions of its id (the low 16 bits of the four bf_hash values,
 * bloomFilter.h:23-30), hashed once on the GPU, n * block * 8 bytes of HBM.
 * The search then reads a row's positions beside its ids instead of hashing
 * them in the hop loop; results are the same either way.  song_index_create
 * builds it when block <= 32.  enabled = 0 frees it (searches hash again),
 * 1 builds it (block <= 32 only).  Not while a search or server on this index
 * runs.  No reference counterpart. */
int song_index_set_visit_table(song_index* index, int32_t enabled);

/* Row addressing of the search.  By default rows are addressed with 32-bit
 * offsets when n < 2^23 and the vectors take under 2 GiB, else with 64-bit
 * ones; enabled = 1 takes the 64-bit path on any index (the same results).
 * No reference counterpart. */
int song_index_set_wide_offsets(song_index* index, int32_t enabled);

/* Floats (or uint64 words for HAMMING) per stored query row on the device:
 * dim rounded up to a multiple of 4 (4 for HAMMING). */
int song_index_row_stride(const song_index* index, int32_t* out);

/* ---------------------------------------------------------------- search */

/* Batched graph search of host queries.  Replaces
 *   gpu_search(const KernelGraph&, const float* queries, nq, k,
 *              N_MULTIQUERY, ef_search, GpuSearchTrace*)   kernalGraph.h:307-382
 *   gpu_search(const HashKernelGraph&, const uint32_t* queries, ...)
 *                                                       hashKernelGraph.h:360-435
 * Per batch: queries H2D, one kernel launch, ids (+ dists, trace) D2H.
 * Results are those of the reference kernel (kg_search_kernel,
 * kernalGraph.h:184-299) for the same graph, seeds, ef and Bloom size.
 *   queries: f32 [nq x dim] or uint64 [nq x 4]
 *   ids_out: [nq x k], best first, -1 padded
 *   dists_out, trace_out, timing_out, batch_out: nullable
 *   batch_out: room for ceil(nq / batch) entries (BatchTiming list). */
int song_search(song_index* index, const void* queries, int32_t nq, int32_t k, int32_t ef,
                const song_search_opts* opts, int32_t* ids_out, float* dists_out,
                song_query_trace* trace_out, song_timing* timing_out,
                song_batch_timing* batch_out);

/* The same search with every buffer already in device memory (queries laid
 * out with song_index_row_stride floats per row); one launch per batch, no
 * copies.  timing_out->kernel_ms is the HIP-event time of the launches on the
 * index stream.  No counterpart in the reference (its gpu_search always
 * copies); used to time the kernel with inputs resident in HBM. */
int song_search_device(song_index* index, const void* d_queries, int32_t nq, int32_t k, int32_t ef,
                       const song_search_opts* opts, int32_t* d_ids, float* d_dists,
                       song_query_trace* d_trace, song_timing* timing_out);

/* song_search_device without any device allocation: the same arguments,
 * kernels and results.  No counterpart in the reference; it is for callers
 * that cannot call hipMalloc (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation).  Its scratch -- the persistent
 * kernel's queue head and the planning pass's workspace -- is one static
 * device array per GPU of 256 + 264 * SONG_STATIC_MAX_PLAN_QUERIES bytes
 * (about 4.3 MB), held until the process ends.  Limit: with learned bit 0 or
 * bit 2 a batch (opts->batch, or nq when 0) holds at most
 * 66 * SONG_STATIC_MAX_PLAN_QUERIES / (n_entry + 2) queries (n_entry = n_seeds,
 * twice that with learned bit 3); a larger one returns SONG_ERR_ARG.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (any
 * index, thread or stream) race, so serialise them. */
#define SONG_STATIC_MAX_PLAN_QUERIES 16384
int song_search_device_static(song_index* index, const void* d_queries, int32_t nq, int32_t k, int32_t ef,
                              const song_search_opts* opts, int32_t* d_ids, float* d_dists,
                              song_query_trace* d_trace, song_timing* timing_out);

/* The reference's CPU search baseline on the host: Graph::search
 * (graph.h:33-115) with the same ids and QueryTrace counters, as bench_cpu
 * runs it (bench_cpu.cpp:444-452), with the queries optionally split over
 * n_threads host threads.  Squared L2 only (summed in index order,
 * distance.h:5-12); the graph's rows have `block` slots (BLOCK_SIZE,
 * config.h:9) and n_seeds evenly spaced entry points (0 = 4, N_MULTIPROBE).
 * This is the CPU method of the bench_cpu driver; the GPU entry points never
 * fall back to it.
 *   ids_out: [nq x k] in Graph::search's own order, WORST first, -1 padded
 *   dists_out, trace_out, per_query_ms_out (each query's wall time),
 *   search_ms_out (wall time of the whole search, argument checks excluded):
 *   nullable. */
int song_cpu_search(const float* vecs, const int32_t* adj, int64_t n, int32_t dim, int32_t block,
                    const float* queries, int32_t nq, int32_t k, int32_t ef, int32_t n_seeds, int32_t n_threads,
                    int32_t* ids_out, float* dists_out, song_query_trace* trace_out, double* per_query_ms_out,
                    double* search_ms_out);

/* Exact k nearest neighbours by brute force on the device, over the index's
 * resident vectors (ground truth for recall).  Replaces the reference's CPU
 * brute force (bench_cpu.cpp:158-175) and numpy compute_gt
 * (run_compare.py:199-213).  Order: distance, then smaller id; k <= 128.
 * kernel_ms_out (nullable): HIP-event time of the two kernels. */
int song_index_bruteforce(song_index* index, const void* queries, int32_t nq, int32_t k,
                          int32_t* ids_out, float* dists_out, double* kernel_ms_out);

/* Device-resident search with a budget per query (mixed k and queue sizes in
 * one launch, as a serving queue holds them).  d_ks / d_efs: int32 [nq] on the
 * device; each ef is raised to its k.  Required, and checked on the host
 * (the budgets are copied back) before anything is launched:
 * 1 <= k <= k_max and max(ef, k) <= max(ef_max, k_max) <= 512; otherwise
 * SONG_ERR_ARG.  Output rows are k_max long, -1 padded.  No counterpart in the
 * reference (one k and ef per gpu_search call, kernalGraph.h:307-309). */
int song_search_device_mixed(song_index* index, const void* d_queries, int32_t nq, const int32_t* d_ks,
                             const int32_t* d_efs, int32_t k_max, int32_t ef_max,
                             const song_search_opts* opts, int32_t* d_ids, float* d_dists,
                             song_query_trace* d_trace, song_timing* timing_out);

/* ---------------------------------------------------------------- learned entry / termination */

/* A trained model for the learned search (no counterpart in the reference).
 * ReLU MLPs with 2 x 64 hidden units, FP32, row-major weights; the query
 * projection z = proj_scale * (proj (q - mu)) (16 values) feeds the first two:
 *   entry: z -> 64 -> 64 -> 64 scores over the 64 pivot nodes; the n_seeds
 *          best are the seeds.  For HAMMING q is the code as a +-1 vector of
 *          256 and mu is unused.
 *   cost:  z -> 64 -> 64 -> 1, the query's predicted standardised log2 hop
 *          count (cost_weights = w1 [64x16], b1 [64], w2 [64x64], b2 [64],
 *          w3 [64], b3 [1]); optional (NULL: no cost order).
 *   stop:  8 per-hop features -> 64 -> 64 -> 1 logit
 *          (stop_weights = w1 [64x8], b1 [64], w2 [64x64], b2 [64], w3 [64]).
 * Entry and cost run in a planning pass before the search (song_plan.hip);
 * the stop features are defined in the search kernel (song_search_impl.h). */
typedef struct song_learned_model {
    int32_t proj_dim;        /* row stride of the index (dim rounded up to 4), or 256 */
    const float* mu;         /* [proj_dim] or NULL (HAMMING) */
    const float* proj;       /* [16 x proj_dim] */
    const float* proj_scale; /* [16] */
    const float* entry_w1;   /* [64 x 16] */
    const float* entry_b1;   /* [64] */
    const float* entry_w2;   /* [64 x 64] */
    const float* entry_b2;   /* [64] */
    const float* entry_w3;   /* [64 x 64] */
    const float* entry_b3;   /* [64] */
    const int32_t* pivots;   /* [64] node ids < n */
    const float* stop_weights; /* [4800] */
    float stop_b3;
    float dist_scale;        /* > 0: distance differences are divided by it */
    const float* cost_weights; /* [5313] or NULL */
} song_learned_model;

/* Copies a model to the index's device (replacing any previous one). */
int song_index_set_model(song_index* index, const song_learned_model* model);

/* Training traces: runs the search with the stop network off and records, at
 * every consulted hop (opts->stop_min_hops / stop_interval), the 8 features
 * and the current k-th best distance: d_trace [nq x max_evals x 9] (device),
 * d_count [nq] rows written, d_final_dk [nq] the k-th best at the end.
 * Entry per opts->learned bit 0 (opts->entry_late_hops must be 0).  k <= 64. */
int song_learn_trace(song_index* index, const void* d_queries, int32_t nq, int32_t k, int32_t ef,
                     const song_search_opts* opts, int32_t max_evals, float* d_trace, int32_t* d_count,
                     float* d_final_dk);

enum {
    SONG_LOSS_DEFAULT = 0, /* logistic on 0/1 labels (out_dim 1), softmax on class ids (out_dim > 1) */
    SONG_LOSS_MSE = 1      /* squared error on real-valued targets (out_dim 1) */
};

/* Trains a 2 x 64 ReLU MLP (in_dim -> 64 -> 64 -> out_dim) on the device with
 * Adam.  Shapes: 8 -> 1 (stop), 16 -> 1 (cost), 16 -> 64 (entry).
 * X: host float [n x in_dim]; targets: host float [n] as `loss` says.
 * weights: host float, w1 [64 x in] b1 [64] w2 [64x64] b2 [64] w3 [out x 64]
 * b3 [out]; read as the starting point and overwritten with the result.
 * loss_out (nullable): [epochs] mean training loss per epoch.
 * Deterministic for a given seed (sample order) and input. */
int song_mlp_train(int32_t device, const float* X, const float* targets, int64_t n, int32_t in_dim,
                   int32_t out_dim, int32_t loss, int32_t epochs, int32_t batch, float lr, uint32_t seed,
                   float* weights, float* loss_out);

/* The same network's forward pass on the device (the inference code the
 * search kernel uses): logits [n x out_dim]. */
int song_mlp_forward(int32_t device, const float* X, int64_t n, int32_t in_dim, int32_t out_dim,
                     const float* weights, float* logits_out);

/* ---------------------------------------------------------------- serving */

/* Continuous-batching server over a resident index (no counterpart in the
 * reference, whose bench_gpu runs fixed batches, kernalGraph.h:343-375).
 * Requests queue from any thread.  Two ways to run them:
 *  - launch slots (resident = 0): a dispatcher admits requests into n_streams
 *    slots of up to max_batch as slots and memory free up; each slot is one
 *    launch, so requests that arrive while others run start in the next free
 *    slot and overlap on the device;
 *  - resident kernel (resident = 1): one persistent grid stays on the device
 *    and serves a ring of max_batch * n_streams request slots in host-visible
 *    memory; each wave takes the next request as soon as its last one is
 *    done, so no request waits for a batch.  The grid leaves by itself after
 *    idle_ms without requests and is relaunched when more arrive.  Supports
 *    the plain and persistent search with the learned stop (learned bit 1).
 * Each request holds row_bytes + 8 + 8*k_max bytes of request state while in
 * flight; admission keeps the total within memory_budget.  A request's result
 * depends only on its query, k, queue size and `search` options. */
typedef struct song_server song_server;

typedef struct song_server_opts {
    int32_t max_batch;        /* requests per launch, 0 = 1024 */
    int32_t n_streams;        /* launch slots (HIP streams), 1..32, 0 = 4; slots only
                                 overlap on the device up to the process's hardware
                                 queues (GPU_MAX_HW_QUEUES, HIP default 4) */
    int64_t memory_budget;    /* bytes of in-flight request state, 0 = 256 MiB */
    int32_t k_max;            /* largest k accepted, 0 = 100 */
    int32_t ef_max;           /* largest queue size accepted (<= 512), 0 = 512 */
    song_search_opts search;  /* options of every launch (mode, learned, ...).  A
                                 learned search serves k <= 64; with k_max above
                                 that, requests with k > 64 fall back to the plain
                                 search (the same options without the learned
                                 ones), in a second launch of the same slot. */
    int32_t resident;         /* 1 = resident kernel (see above), 0 = launch slots */
    int32_t idle_ms;          /* resident kernel: leave after this long idle, 0 = 20 */
} song_server_opts;

typedef struct song_server_stats {
    int64_t submitted;
    int64_t completed;
    int64_t in_flight;
    int64_t launches;
    int64_t deferred;         /* admission passes that found the budget full */
    int64_t peak_bytes;       /* largest in-flight request state seen */
    int64_t memory_budget;
    int32_t slot_capacity;    /* requests per launch (ring slots when resident) after the budget */
    int32_t pad_;
} song_server_stats;

int song_server_create(song_index* index, const song_server_opts* opts, song_server** out);
/* query: one row in the index's device row layout (song_index_row_stride). */
int song_server_submit(song_server* server, const void* query, int32_t k, int32_t ef, int64_t* ticket);
/* 1 = done (results copied, ticket released), 0 = pending, < 0 = error.
 * latency_ms: completion time minus arrival. */
int song_server_poll(song_server* server, int64_t ticket, int32_t* ids_out, float* dists_out,
                     double* latency_ms);
int song_server_wait(song_server* server, int64_t ticket, int32_t timeout_ms, int32_t* ids_out,
                     float* dists_out, double* latency_ms);
int song_server_stats_get(song_server* server, song_server_stats* out);
void song_server_destroy(song_server* server);
/* Load generator: submits nq requests at Poisson arrival times (rate_qps,
 * seeded), waits for all of them and reports each latency from its scheduled
 * arrival; ids_out / dists_out: [nq x k_max], -1 / inf padded; wall_ms_out:
 * first arrival to last completion. */
int song_server_run_poisson(song_server* server, const void* queries, int32_t nq, const int32_t* ks,
                            const int32_t* efs, double rate_qps, uint64_t seed, int32_t* ids_out,
                            float* dists_out, double* latency_ms_out, double* wall_ms_out);

/* ---------------------------------------------------------------- device memory helpers */

/* Plain device buffers for callers without a GPU runtime of their own
 * (no counterpart in the reference, whose gpu_search allocates per call,
 * kernalGraph.h:326-330). */
int song_dev_alloc(int32_t device, int64_t bytes, void** out);
int song_dev_free(void* p);
int song_memcpy_h2d(void* dst, const void* src, int64_t bytes);
/* Device-to-device copy by a plain streaming kernel (16 B per lane), so that
 * hardware byte counters can be calibrated against a known traffic. */
int song_dev_copy(void* dst, const void* src, int64_t bytes, double* kernel_ms_out);
/* Gathers n_ids rows of row_bytes (a multiple of 16) from the device buffer
 * src [src_rows rows] into dst, dst row r = src row ids[r] (ids: a host array,
 * each in [0, src_rows)), by a kernel reading 16 B per lane: a known traffic
 * with the search's access pattern, for calibrating hardware byte counters. */
int song_dev_gather(void* dst, const void* src, int64_t src_rows, int32_t row_bytes, const int32_t* ids,
                    int64_t n_ids, double* kernel_ms_out);
int song_memcpy_d2h(void* dst, const void* src, int64_t bytes);
/* Copies nq host queries into a new device buffer in the index's row layout. */
int song_queries_to_device(song_index* index, const void* queries, int32_t nq, void** d_out);

#ifdef __cplusplus
}
#endif

#endif /* SONG_H */
