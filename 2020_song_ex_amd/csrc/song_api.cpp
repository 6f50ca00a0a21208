// song_api.cpp -- the C ABI of libsong.so (declared in include/song.h):
// index residency, the batched search driver, graph files and device-memory
// helpers.  The reference's counterpart is the host half of kernalGraph.h /
// hashKernelGraph.h (upload_to_gpu, gpu_search).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/song.h"
#include "nsw_build.h"
#include "song_gt.h"
#include "song_kernels.h"
#include "song_server.h"
#include "song_train.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return SONG_ERR_HIP;
}

#define HIP_TRY(expr, what)                          \
    do {                                             \
        hipError_t e_ = (expr);                      \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

double ms_since(std::chrono::steady_clock::time_point a) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
}

bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

// True when the first `cols` floats of each of `rows` rows (row pitch
// `pitch` floats) are all finite.  NaN distances would order as keys above
// every finite distance, where the reference's float scans treat them by
// position (a NaN never compares smaller, gpuHeap.h:28,41), so non-finite
// vectors and queries are refused instead of searched.
bool rows_finite(const float* p, int64_t rows, int64_t cols, int64_t pitch) {
    for (int64_t r = 0; r < rows; r++) {
        const float* row = p + r * pitch;
        for (int64_t c = 0; c < cols; c++)
            if (!std::isfinite(row[c])) return false;
    }
    return true;
}

int ilog2(int64_t v) {
    int r = 0;
    while ((int64_t(1) << r) < v) r++;
    return r;
}

}  // namespace

struct song_index {
    int device = 0;
    hipStream_t stream = nullptr;
    int metric = 0;
    int64_t n = 0;
    int dim = 0;      // floats per vector (256 for Hamming bits)
    int dstride = 0;  // floats (or u64) per stored row
    int block = 0;
    void* d_data = nullptr;
    int32_t* d_adj = nullptr;
    uint32_t* d_vpos = nullptr;  // visited-filter positions per adjacency slot, or null
    bool wide_offsets = false;   // address rows with 64-bit offsets even where 32-bit ones do
    // grow-only workspace for host-buffer searches
    void* d_q = nullptr;
    int32_t* d_ids = nullptr;
    float* d_dists = nullptr;
    int32_t* d_trace = nullptr;
    int64_t cap_q = 0, cap_ids = 0, cap_trace = 0;
    int32_t* d_counter = nullptr;  // persistent-kernel queue head
    float* d_model = nullptr;      // learned model arrays (one allocation)
    size_t model_floats = 0;
    song::PlanArgs plan{};         // query-side networks: device pointers into d_model
    song::LearnedArgs model{};     // stop network: device pointers into d_model
    bool has_model = false;
    bool has_cost = false;
    int32_t* d_plan = nullptr;     // planning-pass workspace: entries | cost | order
    int64_t cap_plan = 0;
    int n_cus = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<char> stage;  // host staging for padded query rows
    std::mutex mu;
};

namespace {

// Plain streaming copy, 16 bytes per lane per step (the reference traffic for
// calibrating hardware byte counters against a known number of bytes).
__global__ void copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, int64_t n16) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// Rows gathered at given ids, 16 bytes per lane: dst row r = src row ids[r].
// Its traffic is known exactly (rows of row_bytes at scattered ids, as the
// search reads them), which calibrates hardware byte counters on the search's
// access pattern rather than on a stream.
__global__ void gather16_kernel(const uint4* __restrict__ src, const int32_t* __restrict__ ids,
                                uint4* __restrict__ dst, int64_t n_rows, int chunks) {
    const int64_t total = n_rows * chunks;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = i / chunks;
        dst[i] = src[(int64_t)ids[r] * chunks + (i - r * chunks)];
    }
}

size_t row_bytes(const song_index* ix) { return (size_t)ix->dstride * (ix->metric == SONG_METRIC_HAMMING ? 8 : 4); }
size_t host_row_bytes(const song_index* ix) {
    return ix->metric == SONG_METRIC_HAMMING ? 32 : (size_t)ix->dim * 4;
}

struct Plan {
    int ef, S, G, qch, bloom_words, bloom_shift, n_seeds, batch, persistent;
    int learned;  // bit 0 entry, bit 1 stop, bit 2 cost-ordered queue, bit 3 keep the reference seeds
    int n_entry;  // seed-row length: n_seeds, or 2 * n_seeds with bit 3
    float stop_threshold;
    int stop_min_hops, stop_interval, stop_max_hops, stop_confirm, result_cap;
    int stop_late_hops, stop_late_interval;
    int entry_late_hops;
    song::LaunchShape shape;
};

int make_plan(const song_index* ix, int32_t nq, int32_t k, int32_t ef, const song_search_opts* o, Plan* p) {
    song_search_opts z;
    std::memset(&z, 0, sizeof z);
    if (!o) o = &z;
    if (nq < 0) return fail(SONG_ERR_ARG, "nq must be >= 0");
    if (k < 1) return fail(SONG_ERR_ARG, "k must be >= 1");
    if (ef < k) ef = k;  // kernalGraph.h:315
    if (ef > 512) return fail(SONG_ERR_ARG, "ef (queue size) must be <= 512 after ef = max(ef, k)");
    p->ef = ef;
    p->S = (ef + 63) / 64;  // set slots per lane: 1..8
    int64_t bits = o->bloom_bits ? o->bloom_bits : 16384;
    if (!is_pow2(bits) || bits < 512 || bits > 65536)
        return fail(SONG_ERR_ARG, "bloom_bits must be a power of two in [512, 65536]");
    p->bloom_words = (int)(bits / 64);
    p->bloom_shift = ilog2(p->bloom_words);
    p->n_seeds = o->n_seeds ? o->n_seeds : 4;
    if (p->n_seeds < 1 || p->n_seeds > 64) return fail(SONG_ERR_ARG, "n_seeds must be in [1, 64]");
    // the learned stop stages its network in each workgroup's LDS: share it over 8 waves
    const int wpb = o->waves_per_block ? o->waves_per_block : ((o->learned & 2) ? 8 : 4);
    if (wpb != 1 && wpb != 2 && wpb != 4 && wpb != 8) return fail(SONG_ERR_ARG, "waves_per_block must be 1, 2, 4 or 8");
    p->shape.waves_per_block = wpb;
    if (o->mode != SONG_MODE_PLAIN && o->mode != SONG_MODE_PERSISTENT) return fail(SONG_ERR_ARG, "unknown mode");
    p->persistent = o->mode == SONG_MODE_PERSISTENT;
    const int wpc = o->waves_per_cu ? o->waves_per_cu : 32;
    if (wpc < 1 || wpc > 32) return fail(SONG_ERR_ARG, "waves_per_cu must be in [1, 32]");
    p->shape.persistent_blocks = std::max(1, (ix->n_cus * wpc + wpb - 1) / wpb);
    p->batch = o->batch > 0 ? o->batch : (nq > 0 ? nq : 1);
    if (o->batch < 0) return fail(SONG_ERR_ARG, "batch must be >= 0");
    if (ix->metric == SONG_METRIC_HAMMING) {
        p->G = 2;
        p->qch = 1;
    } else {
        const int d4 = ix->dstride / 4;
        int G = o->lanes_per_row;
        if (G) {
            if (!is_pow2(G) || G < 2 || G > 64) return fail(SONG_ERR_ARG, "lanes_per_row must be a power of two in [2, 64]");
        } else {
            // two 16-byte chunks per lane (fewest registers per row in flight)
            G = 2;
            while (G < 64 && G * 2 < d4) G *= 2;
        }
        int qch = 2;
        while (G * qch < d4) qch *= 2;
        if (qch > 4 && G != 64) return fail(SONG_ERR_ARG, "lanes_per_row too small for this dim");
        if (qch > 16) return fail(SONG_ERR_ARG, "dim > 4096 is not supported");
        p->G = G;
        p->qch = qch;
    }
    p->learned = o->learned;
    p->stop_threshold = o->stop_threshold;
    p->stop_min_hops = o->stop_min_hops;
    p->stop_interval = o->stop_interval ? o->stop_interval : 1;
    p->stop_max_hops = o->stop_max_hops;
    if (o->stop_max_hops < 0) return fail(SONG_ERR_ARG, "stop_max_hops must be >= 0");
    p->stop_confirm = o->stop_confirm ? o->stop_confirm : 1;
    if (o->stop_confirm < 0 || o->stop_confirm > 64) return fail(SONG_ERR_ARG, "stop_confirm must be in [0, 64]");
    p->result_cap = o->result_cap;
    if (o->result_cap < 0) return fail(SONG_ERR_ARG, "result_cap must be >= 0");
    if (o->result_cap > 0 && o->result_cap < ef) {
        if (!(o->learned & 2)) return fail(SONG_ERR_ARG, "result_cap below the queue size needs the learned stop (bit 1)");
        if (o->stop_max_hops <= 0) return fail(SONG_ERR_ARG, "result_cap below the queue size needs stop_max_hops > 0");
        if (o->result_cap < k) return fail(SONG_ERR_ARG, "result_cap must be >= k");
    }
    if (o->learned < 0 || o->learned > 15) return fail(SONG_ERR_ARG, "learned must be 0..15");
    if ((o->learned & 8) && !(o->learned & 1)) return fail(SONG_ERR_ARG, "learned bit 3 needs the entry network (bit 0)");
    p->n_entry = (o->learned & 8) ? 2 * p->n_seeds : p->n_seeds;
    if (p->n_entry > 64) return fail(SONG_ERR_ARG, "with learned bit 3, n_seeds must be <= 32");
    if (p->stop_interval < 1 || p->stop_min_hops < 0) return fail(SONG_ERR_ARG, "stop_interval >= 1, stop_min_hops >= 0");
    p->stop_late_hops = o->stop_late_hops;
    p->stop_late_interval = o->stop_late_interval;
    if (o->stop_late_hops < 0) return fail(SONG_ERR_ARG, "stop_late_hops must be >= 0");
    if (o->stop_late_hops > 0 && o->stop_late_interval < 1)
        return fail(SONG_ERR_ARG, "stop_late_interval must be >= 1 when stop_late_hops > 0");
    p->entry_late_hops = o->entry_late_hops;
    if (o->entry_late_hops < 0) return fail(SONG_ERR_ARG, "entry_late_hops must be >= 0");
    if (o->entry_late_hops > 0 && ((o->learned & 3) != 3 || (o->learned & 8)))
        return fail(SONG_ERR_ARG, "entry_late_hops needs learned bits 0 and 1 without bit 3");
    if (o->entry_late_hops > 0 && p->S != 1)
        return fail(SONG_ERR_ARG, "entry_late_hops needs a queue of at most 64 after ef = max(ef, k)");
    if (o->learned) {
        if (!ix->has_model) return fail(SONG_ERR_STATE, "learned search needs a model (song_index_set_model)");
        if (k > 64) return fail(SONG_ERR_ARG, "learned search supports k <= 64");
        if (p->qch > 4) return fail(SONG_ERR_ARG, "learned search supports dim <= 1024");
        if ((o->learned & 4) && !ix->has_cost) return fail(SONG_ERR_STATE, "the model has no cost network");
    }
    return SONG_OK;
}

song::SearchArgs make_args(const song_index* ix, const Plan& p, const void* dq, int32_t nq, int32_t k,
                           int32_t* ids, float* dists, int32_t* trace) {
    song::SearchArgs a = {};
    a.data = ix->d_data;
    a.adj = ix->d_adj;
    a.vpos = ix->d_vpos;
    a.queries = dq;
    a.ids_out = ids;
    a.dists_out = dists;
    a.trace_out = trace;
    a.entries = nullptr;
    a.n = ix->n;
    a.dstride = ix->dstride;
    a.off32 = !ix->wide_offsets && ix->n < (int64_t(1) << 23) && ix->n * (int64_t)row_bytes(ix) < (int64_t(1) << 31)
                  ? 1 : 0;
    a.block = ix->block;
    a.k = k;
    a.ef = p.ef;
    a.nq = nq;
    a.bloom_words = p.bloom_words;
    a.bloom_shift = p.bloom_shift;
    a.n_entry = p.n_entry;
    a.ks = nullptr;
    a.efs = nullptr;
    a.k_stride = k;
    a.queue_head = p.persistent ? ix->d_counter : nullptr;
    a.order = nullptr;
    a.resident = nullptr;
    a.learned = (p.learned & 2) ? 1 : 0;
    if (a.learned) {
        a.L = ix->model;
        a.L.use_stop = 1;
        a.L.stop_logit = p.stop_threshold;
        a.L.min_hops = p.stop_min_hops;
        a.L.interval = p.stop_interval;
        a.L.late_hops = p.stop_late_hops;
        a.L.late_interval = p.stop_late_interval;
        a.L.max_hops = p.stop_max_hops;
        a.L.confirm = p.stop_confirm;
        a.L.result_cap = p.result_cap;
        a.L.entry_late_hops = p.entry_late_hops;
        a.L.trace = nullptr;
        a.L.trace_count = nullptr;
        a.L.final_dk = nullptr;
        a.L.max_evals = 0;
    }
    return a;
}

int ensure_ws(song_index* ix, int64_t nq, int32_t k, bool dists, bool trace, double* alloc_ms) {
    auto t0 = std::chrono::steady_clock::now();
    const int64_t need_q = nq * (int64_t)row_bytes(ix);
    if (need_q > ix->cap_q) {
        if (ix->d_q) (void)hipFree(ix->d_q);
        ix->d_q = nullptr;
        ix->cap_q = 0;
        HIP_TRY(hipMalloc(&ix->d_q, need_q), "hipMalloc queries");
        ix->cap_q = need_q;
    }
    const int64_t need_ids = nq * (int64_t)k;
    if (need_ids > ix->cap_ids) {
        if (ix->d_ids) (void)hipFree(ix->d_ids);
        if (ix->d_dists) (void)hipFree(ix->d_dists);
        ix->d_ids = nullptr;
        ix->d_dists = nullptr;
        ix->cap_ids = 0;
        HIP_TRY(hipMalloc(&ix->d_ids, need_ids * 4), "hipMalloc ids");
        HIP_TRY(hipMalloc(&ix->d_dists, need_ids * 4), "hipMalloc dists");
        ix->cap_ids = need_ids;
    }
    if (trace && nq > ix->cap_trace) {
        if (ix->d_trace) (void)hipFree(ix->d_trace);
        ix->d_trace = nullptr;
        ix->cap_trace = 0;
        HIP_TRY(hipMalloc(&ix->d_trace, nq * 12 * 4), "hipMalloc trace");
        ix->cap_trace = nq;
    }
    (void)dists;
    if (alloc_ms) *alloc_ms = ms_since(t0);
    return SONG_OK;
}

// Pads host rows (dim floats) to the device row stride.
const void* stage_rows(song_index* ix, const void* src, int64_t rows) {
    const size_t hb = host_row_bytes(ix), db = row_bytes(ix);
    if (hb == db) return src;
    ix->stage.assign((size_t)rows * db, 0);
    for (int64_t r = 0; r < rows; r++) std::memcpy(ix->stage.data() + r * db, (const char*)src + r * hb, hb);
    return ix->stage.data();
}

// int32 words of planning workspace for nq queries: entries [nq x n_seeds],
// cost [nq], order [nq].
int64_t plan_words(int64_t nq, int n_seeds) { return nq * (n_seeds + 2); }

bool needs_plan(const Plan& p) { return (p.learned & 5) != 0; }

int ensure_plan_ws(song_index* ix, int64_t nq, int n_seeds) {
    const int64_t need = plan_words(nq, n_seeds);
    if (need <= ix->cap_plan) return SONG_OK;
    if (ix->d_plan) (void)hipFree(ix->d_plan);
    ix->d_plan = nullptr;
    ix->cap_plan = 0;
    HIP_TRY(hipMalloc(&ix->d_plan, need * 4), "hipMalloc planning workspace");
    ix->cap_plan = need;
    return SONG_OK;
}

// The planning pass of a learned launch (entry seeds and / or the cost
// order) into ws; points the launch at its outputs.
hipError_t run_plan(const song_index* ix, const Plan& p, song::SearchArgs& a, int32_t* ws, hipStream_t st) {
    if (!needs_plan(p) || a.nq <= 0) return hipSuccess;
    song::PlanArgs P = ix->plan;
    P.queries = a.queries;
    P.nq = a.nq;
    P.n_entry = p.n_entry;
    P.n_ref = (p.learned & 8) ? p.n_seeds : 0;
    P.entries = (p.learned & 1) ? ws : nullptr;
    P.cost = (p.learned & 4) ? (float*)(ws + (size_t)a.nq * p.n_entry) : nullptr;
    int32_t* order = (p.learned & 4) ? ws + (size_t)a.nq * (p.n_entry + 1) : nullptr;
    hipError_t e = song::launch_plan(P, ix->metric, order, st);
    a.entries = P.entries;
    a.order = order;
    return e;
}

}  // namespace

namespace song {

void set_last_error(const std::string& msg) { g_err = msg; }

void index_info(const song_index* ix, int* metric, int* device) {
    *metric = ix->metric;
    *device = ix->device;
}

int launch_mixed_on_stream(song_index* ix, const void* d_q, int nq, const int32_t* d_ks, const int32_t* d_efs,
                           int k_stride, int ef_max, const song_search_opts* opts, int32_t* d_ids, float* d_dists,
                           int32_t* d_counter, int32_t* d_plan_ws, hipStream_t stream, int k_limit) {
    Plan p;
    if (k_limit <= 0 || k_limit > k_stride) k_limit = k_stride;
    int rc = make_plan(ix, nq, k_limit, ef_max, opts, &p);
    if (rc) return rc;
    if (nq == 0) return SONG_OK;
    SearchArgs a = make_args(ix, p, d_q, nq, k_stride, d_ids, d_dists, nullptr);
    a.ks = d_ks;
    a.efs = d_efs;
    a.queue_head = p.persistent ? d_counter : nullptr;
    if (needs_plan(p)) {
        if (!d_plan_ws) return fail(SONG_ERR_ARG, "learned entry / order needs a planning workspace");
        hipError_t e = run_plan(ix, p, a, d_plan_ws, stream);
        if (e != hipSuccess) return hip_fail(e, "planning pass");
    }
    hipError_t e = launch_search(a, ix->metric, p.S, p.G, p.qch, p.shape, stream);
    return e == hipSuccess ? SONG_OK : hip_fail(e, "search kernel launch");
}

int64_t plan_ws_words(int64_t nq) { return plan_words(nq, 64); }

int check_search_opts(const song_index* ix, int32_t k_max, int32_t ef_max, const song_search_opts* opts) {
    Plan p;
    return make_plan(ix, 1, k_max, ef_max, opts, &p);
}

int launch_resident_on_stream(song_index* ix, const void* d_q, int ring, const int32_t* d_ks, const int32_t* d_efs,
                              int k_stride, int ef_max, const song_search_opts* opts, int32_t* d_ids, float* d_dists,
                              const void* d_queue, hipStream_t stream) {
    return fail(SONG_ERR_ARG, "the resident serving kernel is disabled in this build");
    song_search_opts o = *opts;
    if (o.learned & 5) return fail(SONG_ERR_ARG, "the resident server supports the learned stop only (bit 1)");
    o.mode = SONG_MODE_PERSISTENT;
    Plan p;
    int rc = make_plan(ix, ring, k_stride, ef_max, &o, &p);
    if (rc) return rc;
    SearchArgs a = make_args(ix, p, d_q, ring, k_stride, d_ids, d_dists, nullptr);
    a.ks = d_ks;
    a.efs = d_efs;
    a.queue_head = nullptr;
    a.resident = (const ResidentQueue*)d_queue;
    hipError_t e = launch_search(a, ix->metric, p.S, p.G, p.qch, p.shape, stream);
    return e == hipSuccess ? SONG_OK : hip_fail(e, "resident kernel launch");
}

}  // namespace song

extern "C" {

int song_abi_version(void) { return SONG_ABI_VERSION; }

const char* song_last_error(void) { return g_err.c_str(); }

int song_device_count(int32_t* out) {
    if (!out) return fail(SONG_ERR_ARG, "null out");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) c = 0;
    *out = c;
    return SONG_OK;
}

int song_nsw_build(const void* vecs, int64_t n, int32_t dim, int32_t metric, int32_t block, int32_t fixed_deg,
                   int32_t search_deg, int32_t ef_construction, int32_t n_threads, int32_t batch_div,
                   int64_t seq_prefix, int32_t exact_order, int32_t verbose, int32_t* adj_out) {
    if (!vecs || !adj_out) return fail(SONG_ERR_ARG, "null buffer");
    if (n < 1 || n > INT32_MAX) return fail(SONG_ERR_ARG, "n must be in [1, 2^31)");
    if (metric < 0 || metric > 2) return fail(SONG_ERR_ARG, "unknown metric");
    if (metric != SONG_METRIC_HAMMING && dim < 1) return fail(SONG_ERR_ARG, "dim must be >= 1");
    if (block < 2 || block > 64) return fail(SONG_ERR_ARG, "block must be in [2, 64]");
    if (fixed_deg < 1 || fixed_deg >= block) return fail(SONG_ERR_ARG, "fixed_deg must be in [1, block)");
    if (search_deg < 1 || search_deg > fixed_deg) return fail(SONG_ERR_ARG, "search_deg must be in [1, fixed_deg]");
    if (ef_construction < 1) return fail(SONG_ERR_ARG, "ef_construction must be >= 1");
    if (batch_div < 0) return fail(SONG_ERR_ARG, "batch_div must be >= 0");
    song::BuildParams p;
    p.n = n;
    p.dim = metric == SONG_METRIC_HAMMING ? 256 : dim;
    p.metric = metric;
    p.block = block;
    p.fixed_deg = fixed_deg;
    p.search_deg = search_deg;
    p.ef_construction = ef_construction;
    p.n_threads = n_threads < 1 ? 1 : n_threads;
    p.batch_div = batch_div;
    p.seq_prefix = seq_prefix < 1 ? 1 : seq_prefix;
    p.exact_order = exact_order;
    p.verbose = verbose;
    try {
        return song::nsw_build(vecs, p, adj_out);
    } catch (const std::bad_alloc&) {
        return fail(SONG_ERR_ALLOC, "host allocation failed in nsw_build");
    }
}

// Graph file: "SONGGRF1", int64 n, int32 block, metric, dim, reserved, then
// int32 adjacency [n x block].
int song_graph_save(const char* path, const int32_t* adj, int64_t n, int32_t block, int32_t metric, int32_t dim) {
    if (!path || !adj || n < 1 || block < 1) return fail(SONG_ERR_ARG, "bad graph_save arguments");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(SONG_ERR_IO, std::string("cannot open ") + path);
    const char magic[8] = {'S', 'O', 'N', 'G', 'G', 'R', 'F', '1'};
    int32_t hdr[4] = {block, metric, dim, 0};
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(&n, 8, 1, f) == 1 && std::fwrite(hdr, 4, 4, f) == 4 &&
              std::fwrite(adj, 4, (size_t)(n * block), f) == (size_t)(n * block);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? SONG_OK : fail(SONG_ERR_IO, std::string("write failed: ") + path);
}

int song_graph_load_header(const char* path, int64_t* n, int32_t* block, int32_t* metric, int32_t* dim) {
    if (!path || !n || !block || !metric || !dim) return fail(SONG_ERR_ARG, "null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(SONG_ERR_IO, std::string("cannot open ") + path);
    char magic[8];
    int32_t hdr[4];
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "SONGGRF1", 8) == 0 &&
              std::fread(n, 8, 1, f) == 1 && std::fread(hdr, 4, 4, f) == 4;
    std::fclose(f);
    if (!ok) return fail(SONG_ERR_IO, std::string("not a SONG graph file: ") + path);
    *block = hdr[0];
    *metric = hdr[1];
    *dim = hdr[2];
    return SONG_OK;
}

int song_graph_load(const char* path, int32_t* adj_out, int64_t n, int32_t block) {
    int64_t fn;
    int32_t fb, fm, fd;
    int rc = song_graph_load_header(path, &fn, &fb, &fm, &fd);
    if (rc) return rc;
    if (!adj_out || fn != n || fb != block) return fail(SONG_ERR_ARG, "graph file shape differs from the request");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(SONG_ERR_IO, std::string("cannot open ") + path);
    bool ok = std::fseek(f, 8 + 8 + 16, SEEK_SET) == 0 &&
              std::fread(adj_out, 4, (size_t)(n * block), f) == (size_t)(n * block);
    std::fclose(f);
    return ok ? SONG_OK : fail(SONG_ERR_IO, std::string("short read: ") + path);
}

int song_index_create(const void* vecs, const int32_t* adj, int64_t n, int32_t dim, int32_t block, int32_t metric,
                      int32_t device, song_index** out) {
    if (!out) return fail(SONG_ERR_ARG, "null out");
    *out = nullptr;
    if (!vecs || !adj) return fail(SONG_ERR_ARG, "null buffer");
    if (n < 1 || n > INT32_MAX) return fail(SONG_ERR_ARG, "n must be in [1, 2^31)");
    if (metric < 0 || metric > 2) return fail(SONG_ERR_ARG, "unknown metric");
    if (block < 1 || block > 64) return fail(SONG_ERR_ARG, "block must be in [1, 64]");
    if (metric == SONG_METRIC_HAMMING) {
        if (dim != 256) return fail(SONG_ERR_ARG, "Hamming codes must have dim = 256 bits");
    } else if (dim < 1 || dim > 4096) {
        return fail(SONG_ERR_ARG, "dim must be in [1, 4096]");
    }
    for (int64_t i = 0; i < n * block; i++) {
        if (adj[i] < -1 || adj[i] >= n) return fail(SONG_ERR_ARG, "adjacency entry out of range at " + std::to_string(i));
    }
    if (metric != SONG_METRIC_HAMMING && !rows_finite((const float*)vecs, n, dim, dim))
        return fail(SONG_ERR_ARG, "vectors must be finite (NaN or Inf found)");
    int nd = 0;
    HIP_TRY(hipGetDeviceCount(&nd), "hipGetDeviceCount");
    if (device < 0 || device >= nd) return fail(SONG_ERR_ARG, "device out of range");
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    song_index* ix = new song_index();
    ix->device = device;
    ix->metric = metric;
    ix->n = n;
    ix->dim = dim;
    ix->dstride = metric == SONG_METRIC_HAMMING ? 4 : (dim + 3) / 4 * 4;
    ix->block = block;
    auto bail = [&](int rc) {
        song_index_destroy(ix);
        return rc;
    };
    hipError_t e = hipStreamCreateWithFlags(&ix->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return bail(hip_fail(e, "hipStreamCreate"));
    for (auto& ev : ix->ev) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) return bail(hip_fail(e, "hipEventCreate"));
    }
    const size_t rb = row_bytes(ix);
    e = hipMalloc(&ix->d_data, (size_t)n * rb);
    if (e != hipSuccess) return bail(hip_fail(e, "hipMalloc vectors"));
    e = hipMalloc(&ix->d_adj, (size_t)n * block * 4);
    if (e != hipSuccess) return bail(hip_fail(e, "hipMalloc adjacency"));
    e = hipMalloc(&ix->d_counter, 64);
    if (e != hipSuccess) return bail(hip_fail(e, "hipMalloc counter"));
    e = hipDeviceGetAttribute(&ix->n_cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return bail(hip_fail(e, "hipDeviceGetAttribute"));
    const size_t hb = host_row_bytes(ix);
    if (hb == rb) {
        e = hipMemcpy(ix->d_data, vecs, (size_t)n * rb, hipMemcpyHostToDevice);
    } else {
        // pad rows in chunks
        const int64_t chunk = 65536;
        std::vector<char> buf((size_t)chunk * rb, 0);
        for (int64_t r0 = 0; r0 < n && e == hipSuccess; r0 += chunk) {
            const int64_t rows = std::min(chunk, n - r0);
            for (int64_t r = 0; r < rows; r++)
                std::memcpy(buf.data() + r * rb, (const char*)vecs + (r0 + r) * hb, hb);
            e = hipMemcpy((char*)ix->d_data + r0 * rb, buf.data(), (size_t)rows * rb, hipMemcpyHostToDevice);
        }
    }
    if (e != hipSuccess) return bail(hip_fail(e, "upload vectors"));
    e = hipMemcpy(ix->d_adj, adj, (size_t)n * block * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) return bail(hip_fail(e, "upload adjacency"));
    if (block <= 32) {
        const int rc = song_index_set_visit_table(ix, 1);
        if (rc != SONG_OK) return bail(rc);
    }
    *out = ix;
    return SONG_OK;
}

int song_index_set_wide_offsets(song_index* ix, int32_t enabled) {
    if (!ix) return fail(SONG_ERR_ARG, "null index");
    std::lock_guard<std::mutex> lk(ix->mu);
    ix->wide_offsets = enabled != 0;
    return SONG_OK;
}

int song_index_set_visit_table(song_index* ix, int32_t enabled) {
    if (!ix) return fail(SONG_ERR_ARG, "null index");
    std::lock_guard<std::mutex> lk(ix->mu);
    HIP_TRY(hipSetDevice(ix->device), "hipSetDevice");
    if (!enabled) {
        if (ix->d_vpos) {
            HIP_TRY(hipStreamSynchronize(ix->stream), "hipStreamSynchronize");
            HIP_TRY(hipFree(ix->d_vpos), "hipFree visit table");
            ix->d_vpos = nullptr;
        }
        return SONG_OK;
    }
    if (ix->block > 32) return fail(SONG_ERR_ARG, "the visit table serves rows of at most 32 slots");
    if (ix->d_vpos) return SONG_OK;
    uint32_t* t = nullptr;
    HIP_TRY(hipMalloc(&t, (size_t)ix->n * 2 * ix->block * 4), "hipMalloc visit table");
    hipError_t e = song::launch_visit_table(ix->d_adj, ix->n, ix->block, t, ix->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ix->stream);
    if (e != hipSuccess) {
        (void)hipFree(t);
        return hip_fail(e, "visit table kernel");
    }
    ix->d_vpos = t;
    return SONG_OK;
}

void song_index_destroy(song_index* ix) {
    if (!ix) return;
    (void)hipSetDevice(ix->device);
    if (ix->stream) (void)hipStreamSynchronize(ix->stream);
    if (ix->d_data) (void)hipFree(ix->d_data);
    if (ix->d_adj) (void)hipFree(ix->d_adj);
    if (ix->d_vpos) (void)hipFree(ix->d_vpos);
    if (ix->d_q) (void)hipFree(ix->d_q);
    if (ix->d_ids) (void)hipFree(ix->d_ids);
    if (ix->d_dists) (void)hipFree(ix->d_dists);
    if (ix->d_trace) (void)hipFree(ix->d_trace);
    if (ix->d_counter) (void)hipFree(ix->d_counter);
    if (ix->d_model) (void)hipFree(ix->d_model);
    if (ix->d_plan) (void)hipFree(ix->d_plan);
    for (auto& ev : ix->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (ix->stream) (void)hipStreamDestroy(ix->stream);
    delete ix;
}

int song_index_row_stride(const song_index* ix, int32_t* out) {
    if (!ix || !out) return fail(SONG_ERR_ARG, "null argument");
    *out = ix->dstride;
    return SONG_OK;
}

int song_search(song_index* ix, const void* queries, int32_t nq, int32_t k, int32_t ef, const song_search_opts* opts,
                int32_t* ids_out, float* dists_out, song_query_trace* trace_out, song_timing* timing_out,
                song_batch_timing* batch_out) {
    if (!ix) return fail(SONG_ERR_ARG, "null index");
    if (nq > 0 && (!queries || !ids_out)) return fail(SONG_ERR_ARG, "null buffer");
    Plan p;
    int rc = make_plan(ix, nq, k, ef, opts, &p);
    if (rc) return rc;
    if (ix->metric != SONG_METRIC_HAMMING && !rows_finite((const float*)queries, nq, ix->dim, ix->dim))
        return fail(SONG_ERR_ARG, "queries must be finite (NaN or Inf found)");
    std::lock_guard<std::mutex> lock(ix->mu);
    HIP_TRY(hipSetDevice(ix->device), "hipSetDevice");
    song_timing tm;
    std::memset(&tm, 0, sizeof tm);
    tm.ef_used = p.ef;
    const int64_t bmax = std::min<int64_t>(p.batch, std::max<int32_t>(nq, 1));
    rc = ensure_ws(ix, bmax, k, dists_out != nullptr, trace_out != nullptr, &tm.alloc_ms);
    if (rc) return rc;
    if (needs_plan(p) && (rc = ensure_plan_ws(ix, bmax, p.n_entry))) return rc;
    const size_t hb = host_row_bytes(ix), rb = row_bytes(ix);
    auto t0 = std::chrono::steady_clock::now();
    int bi = 0;
    for (int64_t off = 0; off < nq; off += p.batch, bi++) {
        const int32_t bs = (int32_t)std::min<int64_t>(p.batch, nq - off);
        const void* src = stage_rows(ix, (const char*)queries + off * hb, bs);
        HIP_TRY(hipEventRecord(ix->ev[0], ix->stream), "hipEventRecord");
        HIP_TRY(hipMemcpyAsync(ix->d_q, src, (size_t)bs * rb, hipMemcpyHostToDevice, ix->stream), "query H2D");
        HIP_TRY(hipEventRecord(ix->ev[1], ix->stream), "hipEventRecord");
        song::SearchArgs a = make_args(ix, p, ix->d_q, bs, k, ix->d_ids, dists_out ? ix->d_dists : nullptr,
                                       trace_out ? ix->d_trace : nullptr);
        HIP_TRY(run_plan(ix, p, a, ix->d_plan, ix->stream), "planning pass");
        HIP_TRY(song::launch_search(a, ix->metric, p.S, p.G, p.qch, p.shape, ix->stream), "search kernel launch");
        HIP_TRY(hipEventRecord(ix->ev[2], ix->stream), "hipEventRecord");
        HIP_TRY(hipMemcpyAsync(ids_out + off * k, ix->d_ids, (size_t)bs * k * 4, hipMemcpyDeviceToHost, ix->stream),
                "ids D2H");
        if (dists_out)
            HIP_TRY(hipMemcpyAsync(dists_out + off * k, ix->d_dists, (size_t)bs * k * 4, hipMemcpyDeviceToHost,
                                   ix->stream),
                    "dists D2H");
        if (trace_out)
            HIP_TRY(hipMemcpyAsync(trace_out + off, ix->d_trace, (size_t)bs * 48, hipMemcpyDeviceToHost, ix->stream),
                    "trace D2H");
        HIP_TRY(hipEventRecord(ix->ev[3], ix->stream), "hipEventRecord");
        HIP_TRY(hipEventSynchronize(ix->ev[3]), "search synchronize");
        float h2d = 0, ker = 0, d2h = 0;
        HIP_TRY(hipEventElapsedTime(&h2d, ix->ev[0], ix->ev[1]), "hipEventElapsedTime");
        HIP_TRY(hipEventElapsedTime(&ker, ix->ev[1], ix->ev[2]), "hipEventElapsedTime");
        HIP_TRY(hipEventElapsedTime(&d2h, ix->ev[2], ix->ev[3]), "hipEventElapsedTime");
        tm.h2d_ms += h2d;
        tm.kernel_ms += ker;
        tm.d2h_ms += d2h;
        if (batch_out) {
            song_batch_timing& b = batch_out[bi];
            b.batch_index = bi;
            b.batch_offset = (int32_t)off;
            b.batch_size = bs;
            b.pad_ = 0;
            b.h2d_ms = h2d;
            b.memset_ms = 0.0;
            b.kernel_ms = ker;
            b.d2h_ms = d2h;
        }
    }
    tm.total_ms = ms_since(t0);
    tm.n_batches = bi;
    if (timing_out) *timing_out = tm;
    return SONG_OK;
}

}  // extern "C"

namespace {

// The batch loop of song_search_device and song_search_device_static over the
// caller's device buffers: `counter` is the persistent kernel's queue head and
// `plan_ws` the planning pass's workspace (large enough for one batch).
// Called with ix->mu held and the index's device current.
int search_device_batches(song_index* ix, const Plan& p, const void* d_queries, int32_t nq, int32_t k,
                          int32_t* d_ids, float* d_dists, song_query_trace* d_trace, int32_t* counter,
                          int32_t* plan_ws, song_timing* timing_out) {
    song_timing tm;
    std::memset(&tm, 0, sizeof tm);
    tm.ef_used = p.ef;
    const size_t rb = row_bytes(ix);
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipEventRecord(ix->ev[1], ix->stream), "hipEventRecord");
    int bi = 0;
    for (int64_t off = 0; off < nq; off += p.batch, bi++) {
        const int32_t bs = (int32_t)std::min<int64_t>(p.batch, nq - off);
        song::SearchArgs a = make_args(ix, p, (const char*)d_queries + off * rb, bs, k, d_ids + off * k,
                                       d_dists ? d_dists + off * k : nullptr,
                                       d_trace ? (int32_t*)(d_trace + off) : nullptr);
        a.queue_head = p.persistent ? counter : nullptr;
        HIP_TRY(run_plan(ix, p, a, plan_ws, ix->stream), "planning pass");
        HIP_TRY(song::launch_search(a, ix->metric, p.S, p.G, p.qch, p.shape, ix->stream), "search kernel launch");
    }
    HIP_TRY(hipEventRecord(ix->ev[2], ix->stream), "hipEventRecord");
    HIP_TRY(hipEventSynchronize(ix->ev[2]), "search synchronize");
    float ker = 0;
    HIP_TRY(hipEventElapsedTime(&ker, ix->ev[1], ix->ev[2]), "hipEventElapsedTime");
    tm.kernel_ms = ker;
    tm.total_ms = ms_since(t0);
    tm.n_batches = bi;
    if (timing_out) *timing_out = tm;
    return SONG_OK;
}

// Scratch of song_search_device_static, one copy per GPU for the life of the
// process: words [0, 64) hold the persistent kernel's queue head (zeroed by a
// kernel on the stream before each launch), the rest the planning pass's
// workspace, entries | cost | order, for one batch (every word it reads is
// written first in the same call).  In host code `scratch` only names the
// symbol; its device address comes from hipGetSymbolAddress, per device.
constexpr int64_t kScratchHeadWords = 64;
constexpr int64_t kScratchPlanWords = (int64_t)SONG_STATIC_MAX_PLAN_QUERIES * (64 + 2);
static __device__ __align__(256) int32_t scratch[kScratchHeadWords + kScratchPlanWords];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch_ptr[kMaxScratchDevices] = {};

// The device address of `scratch` on `device`, resolved once per device with
// that device current (each GPU holds its own copy at its own address).
int scratch_on_device(int device, int32_t** out) {
    if (device < 0 || device >= kMaxScratchDevices) return fail(SONG_ERR_ARG, "device id out of range for scratch");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        HIP_TRY(hipSetDevice(device), "hipSetDevice");
        void* ptr = nullptr;
        HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(SONG_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[device] = (int32_t*)ptr;
    }
    *out = g_scratch_ptr[device];
    return SONG_OK;
}

}  // namespace

extern "C" {

int song_search_device(song_index* ix, const void* d_queries, int32_t nq, int32_t k, int32_t ef,
                       const song_search_opts* opts, int32_t* d_ids, float* d_dists, song_query_trace* d_trace,
                       song_timing* timing_out) {
    if (!ix) return fail(SONG_ERR_ARG, "null index");
    if (nq > 0 && (!d_queries || !d_ids)) return fail(SONG_ERR_ARG, "null buffer");
    Plan p;
    int rc = make_plan(ix, nq, k, ef, opts, &p);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(ix->mu);
    HIP_TRY(hipSetDevice(ix->device), "hipSetDevice");
    if (needs_plan(p) && (rc = ensure_plan_ws(ix, std::min<int64_t>(p.batch, std::max<int32_t>(nq, 1)), p.n_entry)))
        return rc;
    return search_device_batches(ix, p, d_queries, nq, k, d_ids, d_dists, d_trace, ix->d_counter, ix->d_plan,
                                 timing_out);
}

int song_search_device_static(song_index* ix, const void* d_queries, int32_t nq, int32_t k, int32_t ef,
                              const song_search_opts* opts, int32_t* d_ids, float* d_dists,
                              song_query_trace* d_trace, song_timing* timing_out) {
    if (!ix) return fail(SONG_ERR_ARG, "null index");
    if (nq > 0 && (!d_queries || !d_ids)) return fail(SONG_ERR_ARG, "null buffer");
    Plan p;
    int rc = make_plan(ix, nq, k, ef, opts, &p);
    if (rc) return rc;
    const int64_t batch = std::min<int64_t>(p.batch, std::max<int32_t>(nq, 1));
    if (needs_plan(p) && plan_words(batch, p.n_entry) > kScratchPlanWords)
        return fail(SONG_ERR_ARG, "song_search_device_static: with the entry network or the cost order a batch "
                                  "holds at most " + std::to_string(kScratchPlanWords / (p.n_entry + 2)) +
                                  " queries for this n_seeds (set opts->batch)");
    std::lock_guard<std::mutex> lock(ix->mu);
    int32_t* s = nullptr;
    if ((rc = scratch_on_device(ix->device, &s))) return rc;
    HIP_TRY(hipSetDevice(ix->device), "hipSetDevice");
    return search_device_batches(ix, p, d_queries, nq, k, d_ids, d_dists, d_trace, s, s + kScratchHeadWords,
                                 timing_out);
}

int song_search_device_mixed(song_index* ix, const void* d_queries, int32_t nq, const int32_t* d_ks,
                             const int32_t* d_efs, int32_t k_max, int32_t ef_max, const song_search_opts* opts,
                             int32_t* d_ids, float* d_dists, song_query_trace* d_trace, song_timing* timing_out) {
    if (!ix) return fail(SONG_ERR_ARG, "null index");
    if (nq > 0 && (!d_queries || !d_ids || !d_ks || !d_efs)) return fail(SONG_ERR_ARG, "null buffer");
    Plan p;
    int rc = make_plan(ix, nq, k_max, ef_max, opts, &p);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(ix->mu);
    HIP_TRY(hipSetDevice(ix->device), "hipSetDevice");
    // every budget is checked on the host before the launch: 1 <= k <= k_max
    // and max(ef, k) <= ef_max (the kernel sizes its sets and rows by them)
    if (nq > 0) {
        std::vector<int32_t> hk((size_t)nq), he((size_t)nq);
        HIP_TRY(hipMemcpyAsync(hk.data(), d_ks, (size_t)nq * 4, hipMemcpyDeviceToHost, ix->stream), "budgets D2H");
        HIP_TRY(hipMemcpyAsync(he.data(), d_efs, (size_t)nq * 4, hipMemcpyDeviceToHost, ix->stream), "budgets D2H");
        HIP_TRY(hipStreamSynchronize(ix->stream), "budgets D2H");
        for (int32_t i = 0; i < nq; i++) {
            if (hk[i] < 1 || hk[i] > k_max)
                return fail(SONG_ERR_ARG, "query " + std::to_string(i) + ": k must be in [1, k_max]");
            if (std::max(he[i], hk[i]) > p.ef)
                return fail(SONG_ERR_ARG, "query " + std::to_string(i) + ": max(ef, k) must be <= ef_max");
        }
    }
    song_timing tm;
    std::memset(&tm, 0, sizeof tm);
    tm.ef_used = p.ef;
    if (needs_plan(p) && (rc = ensure_plan_ws(ix, std::max<int32_t>(nq, 1), p.n_entry))) return rc;
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipEventRecord(ix->ev[1], ix->stream), "hipEventRecord");
    if (nq > 0) {
        song::SearchArgs a = make_args(ix, p, d_queries, nq, k_max, d_ids, d_dists, (int32_t*)d_trace);
        a.ks = d_ks;
        a.efs = d_efs;
        HIP_TRY(run_plan(ix, p, a, ix->d_plan, ix->stream), "planning pass");
        HIP_TRY(song::launch_search(a, ix->metric, p.S, p.G, p.qch, p.shape, ix->stream), "search kernel launch");
    }
    HIP_TRY(hipEventRecord(ix->ev[2], ix->stream), "hipEventRecord");
    HIP_TRY(hipEventSynchronize(ix->ev[2]), "search synchronize");
    float ker = 0;
    HIP_TRY(hipEventElapsedTime(&ker, ix->ev[1], ix->ev[2]), "hipEventElapsedTime");
    tm.kernel_ms = ker;
    tm.total_ms = ms_since(t0);
    tm.n_batches = nq > 0 ? 1 : 0;
    if (timing_out) *timing_out = tm;
    return SONG_OK;
}

int song_index_set_model(song_index* ix, const song_learned_model* m) {
    if (!ix || !m) return fail(SONG_ERR_ARG, "null argument");
    const int pd = ix->metric == SONG_METRIC_HAMMING ? 256 : ix->dstride;
    if (m->proj_dim != pd) return fail(SONG_ERR_ARG, "proj_dim must be the index row stride (256 for Hamming)");
    if (!m->proj || !m->proj_scale || !m->entry_w1 || !m->entry_b1 || !m->entry_w2 || !m->entry_b2 ||
        !m->entry_w3 || !m->entry_b3 || !m->pivots || !m->stop_weights || (ix->metric != SONG_METRIC_HAMMING && !m->mu))
        return fail(SONG_ERR_ARG, "null model array");
    if (!(m->dist_scale > 0.f)) return fail(SONG_ERR_ARG, "dist_scale must be > 0");
    for (int c = 0; c < song::LM_C; c++)
        if (m->pivots[c] < 0 || m->pivots[c] >= ix->n) return fail(SONG_ERR_ARG, "pivot id out of range");
    using song::LM_C;
    using song::LM_H;
    using song::LM_P;
    // one allocation: mu | proj | scale | e_w1 | e_b1 | e_w2 | e_b2 | e_w3 | e_b3 | pivots | stop | cost
    const size_t sizes[] = {(size_t)pd, (size_t)LM_P * pd, LM_P, LM_H * LM_P, LM_H, LM_H * LM_H, LM_H,
                            (size_t)LM_C * LM_H, LM_C, LM_C, song::LM_STOP_FLOATS, song::LM_COST_FLOATS};
    std::vector<size_t> offs;
    size_t total = 0;
    for (size_t sz : sizes) {
        offs.push_back(total);
        total += (sz + 3) / 4 * 4;  // keep every array 16-byte aligned
    }
    std::vector<float> host(total, 0.f);
    auto put = [&](int i, const float* src) {
        if (src) std::memcpy(host.data() + offs[i], src, sizes[i] * 4);
    };
    put(0, m->mu);
    put(1, m->proj);
    put(2, m->proj_scale);
    put(3, m->entry_w1);
    put(4, m->entry_b1);
    put(5, m->entry_w2);
    put(6, m->entry_b2);
    put(7, m->entry_w3);
    put(8, m->entry_b3);
    std::memcpy(host.data() + offs[9], m->pivots, LM_C * 4);
    song::stop_weights_to_kernel_layout(m->stop_weights, host.data() + offs[10]);
    put(11, m->cost_weights);
    std::lock_guard<std::mutex> lock(ix->mu);
    HIP_TRY(hipSetDevice(ix->device), "hipSetDevice");
    if (ix->d_model && ix->model_floats < total) {
        HIP_TRY(hipStreamSynchronize(ix->stream), "hipStreamSynchronize");
        (void)hipFree(ix->d_model);
        ix->d_model = nullptr;
        ix->model_floats = 0;
        ix->has_model = ix->has_cost = false;
    }
    if (!ix->d_model) {
        HIP_TRY(hipMalloc(&ix->d_model, total * 4), "hipMalloc model");
        ix->model_floats = total;
    }
    HIP_TRY(hipMemcpy(ix->d_model, host.data(), total * 4, hipMemcpyHostToDevice), "model upload");
    float* b = ix->d_model;
    song::PlanArgs& P = ix->plan;
    P = song::PlanArgs{};
    P.dstride = ix->dstride;
    P.n = ix->n;
    P.mu = b + offs[0];
    P.proj = b + offs[1];
    P.proj_scale = b + offs[2];
    P.e_w1 = b + offs[3];
    P.e_b1 = b + offs[4];
    P.e_w2 = b + offs[5];
    P.e_b2 = b + offs[6];
    P.e_w3 = b + offs[7];
    P.e_b3 = b + offs[8];
    P.pivots = (const int32_t*)(b + offs[9]);
    P.c_weights = m->cost_weights ? b + offs[11] : nullptr;
    song::LearnedArgs& L = ix->model;
    L = song::LearnedArgs{};
    L.s_weights = b + offs[10];
    L.s_b3 = m->stop_b3;
    L.dist_scale = m->dist_scale;
    ix->has_model = true;
    ix->has_cost = m->cost_weights != nullptr;
    return SONG_OK;
}

int song_learn_trace(song_index* ix, const void* d_queries, int32_t nq, int32_t k, int32_t ef,
                     const song_search_opts* opts, int32_t max_evals, float* d_trace, int32_t* d_count,
                     float* d_final_dk) {
    if (!ix) return fail(SONG_ERR_ARG, "null index");
    if (nq > 0 && (!d_queries || !d_trace || !d_count || !d_final_dk)) return fail(SONG_ERR_ARG, "null buffer");
    if (max_evals < 1) return fail(SONG_ERR_ARG, "max_evals must be >= 1");
    song_search_opts o;
    if (opts) o = *opts;
    else std::memset(&o, 0, sizeof o);
    o.learned = (o.learned & 1) | 2;  // trace mode rides on the stop path
    if (o.entry_late_hops != 0) return fail(SONG_ERR_ARG, "song_learn_trace records searches without entry_late_hops");
    Plan p;
    int rc = make_plan(ix, nq, k, ef, &o, &p);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(ix->mu);
    HIP_TRY(hipSetDevice(ix->device), "hipSetDevice");
    if (needs_plan(p) && (rc = ensure_plan_ws(ix, std::max<int32_t>(nq, 1), p.n_entry))) return rc;
    int32_t* d_ids = nullptr;
    HIP_TRY(hipMalloc(&d_ids, (size_t)std::max(nq, 1) * k * 4), "hipMalloc ids");
    hipError_t e = hipSuccess;
    if (nq > 0) {
        song::SearchArgs a = make_args(ix, p, d_queries, nq, k, d_ids, nullptr, nullptr);
        a.L.trace = d_trace;
        a.L.trace_count = d_count;
        a.L.final_dk = d_final_dk;
        a.L.max_evals = max_evals;
        e = run_plan(ix, p, a, ix->d_plan, ix->stream);
        if (e == hipSuccess) e = song::launch_search(a, ix->metric, p.S, p.G, p.qch, p.shape, ix->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ix->stream);
    }
    (void)hipFree(d_ids);
    if (e != hipSuccess) return hip_fail(e, "trace search");
    return SONG_OK;
}

int song_mlp_train(int32_t device, const float* X, const float* targets, int64_t n, int32_t in_dim, int32_t out_dim,
                   int32_t loss, int32_t epochs, int32_t batch, float lr, uint32_t seed, float* weights,
                   float* loss_out) {
    if (!X || !targets || !weights || n < 1 || epochs < 1 || batch < 1 || !(lr > 0.f))
        return fail(SONG_ERR_ARG, "bad training arguments");
    if (!((in_dim == 8 && out_dim == 1) || (in_dim == 16 && out_dim == 1) || (in_dim == 16 && out_dim == 64)))
        return fail(SONG_ERR_ARG, "supported shapes: 8 -> 1 (stop), 16 -> 1 (cost), 16 -> 64 (entry)");
    if (loss != SONG_LOSS_DEFAULT && loss != SONG_LOSS_MSE) return fail(SONG_ERR_ARG, "unknown loss");
    if (loss == SONG_LOSS_MSE && out_dim != 1) return fail(SONG_ERR_ARG, "the squared-error loss needs out_dim == 1");
    if (out_dim > 1)
        for (int64_t i = 0; i < n; i++)
            if (!(targets[i] >= 0.f && targets[i] < (float)out_dim)) return fail(SONG_ERR_ARG, "class id out of range");
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    const int total = song::mlp_param_count(in_dim, out_dim);
    const int per_wave = 8;
    const int nwaves_max = (batch + per_wave - 1) / per_wave;
    float *dX = nullptr, *dT = nullptr, *dW = nullptr, *dM = nullptr, *dV = nullptr, *dP = nullptr, *dL = nullptr;
    int32_t* dO = nullptr;
    auto cleanup = [&]() {
        for (void* q : {(void*)dX, (void*)dT, (void*)dW, (void*)dM, (void*)dV, (void*)dP, (void*)dL, (void*)dO})
            if (q) (void)hipFree(q);
    };
    hipError_t e = hipMalloc(&dX, (size_t)n * in_dim * 4);
    if (e == hipSuccess) e = hipMalloc(&dT, (size_t)n * 4);
    if (e == hipSuccess) e = hipMalloc(&dW, (size_t)total * 4);
    if (e == hipSuccess) e = hipMalloc(&dM, (size_t)total * 4);
    if (e == hipSuccess) e = hipMalloc(&dV, (size_t)total * 4);
    if (e == hipSuccess) e = hipMalloc(&dP, (size_t)nwaves_max * (total + 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&dL, (size_t)epochs * 4);
    if (e == hipSuccess) e = hipMalloc(&dO, (size_t)n * 4);
    if (e == hipSuccess) e = hipMemcpy(dX, X, (size_t)n * in_dim * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dT, targets, (size_t)n * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dW, weights, (size_t)total * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(dM, 0, (size_t)total * 4);
    if (e == hipSuccess) e = hipMemset(dV, 0, (size_t)total * 4);
    if (e == hipSuccess) e = hipMemset(dL, 0, (size_t)epochs * 4);
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(e, "training setup");
    }
    std::vector<int32_t> order((size_t)n);
    uint64_t rng = 0x9E3779B97F4A7C15ull ^ seed;
    auto next = [&]() {  // splitmix64
        uint64_t z = (rng += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    int64_t step = 0;
    for (int ep = 0; ep < epochs && e == hipSuccess; ep++) {
        for (int64_t i = 0; i < n; i++) order[i] = (int32_t)i;
        for (int64_t i = n - 1; i > 0; i--) std::swap(order[i], order[next() % (uint64_t)(i + 1)]);
        e = hipMemcpy(dO, order.data(), (size_t)n * 4, hipMemcpyHostToDevice);
        for (int64_t b0 = 0; b0 < n && e == hipSuccess; b0 += batch) {
            const int nb = (int)std::min<int64_t>(batch, n - b0);
            const int nw = (nb + per_wave - 1) / per_wave;
            step++;
            const float bc1 = 1.f - std::pow(0.9f, (float)step), bc2 = 1.f - std::pow(0.999f, (float)step);
            e = song::launch_mlp_grad(in_dim, out_dim, loss == SONG_LOSS_MSE, dX, dT, dO, b0, nb, per_wave, dW, dP,
                                      nw, 0);
            if (e == hipSuccess) e = song::launch_mlp_adam(dW, dM, dV, dP, nw, total, nb, lr, bc1, bc2, dL + ep, 0);
        }
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(weights, dW, (size_t)total * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess && loss_out) {
        e = hipMemcpy(loss_out, dL, (size_t)epochs * 4, hipMemcpyDeviceToHost);
        for (int ep = 0; ep < epochs; ep++) loss_out[ep] /= (float)n;
    }
    cleanup();
    if (e != hipSuccess) return hip_fail(e, "training");
    return SONG_OK;
}

int song_mlp_forward(int32_t device, const float* X, int64_t n, int32_t in_dim, int32_t out_dim,
                     const float* weights, float* logits_out) {
    if (!X || !weights || !logits_out || n < 0) return fail(SONG_ERR_ARG, "bad arguments");
    if (!((in_dim == 8 && out_dim == 1) || (in_dim == 16 && out_dim == 1) || (in_dim == 16 && out_dim == 64)))
        return fail(SONG_ERR_ARG, "supported shapes: 8 -> 1 (stop), 16 -> 1 (cost), 16 -> 64 (entry)");
    if (n == 0) return SONG_OK;
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    const int total = song::mlp_param_count(in_dim, out_dim);
    float *dX = nullptr, *dW = nullptr, *dY = nullptr;
    hipError_t e = hipMalloc(&dX, (size_t)n * in_dim * 4);
    if (e == hipSuccess) e = hipMalloc(&dW, (size_t)(total + 4) * 4);
    if (e == hipSuccess) e = hipMalloc(&dY, (size_t)n * out_dim * 4);
    if (e == hipSuccess) e = hipMemcpy(dX, X, (size_t)n * in_dim * 4, hipMemcpyHostToDevice);
    std::vector<float> wk(weights, weights + total);
    if (in_dim == 8) song::stop_weights_to_kernel_layout(weights, wk.data());
    if (e == hipSuccess) e = hipMemcpy(dW, wk.data(), (size_t)total * 4, hipMemcpyHostToDevice);
    song::LearnedArgs L{};
    song::PlanArgs P{};
    int which = 0;
    if (in_dim == 8) {
        L.s_weights = dW;
        L.s_b3 = weights[song::LM_STOP_FLOATS];
    } else if (out_dim == 1) {
        which = 2;
        P.c_weights = dW;
    } else {
        const int H = song::LM_H;
        which = 1;
        P.e_w1 = dW;
        P.e_b1 = dW + H * 16;
        P.e_w2 = P.e_b1 + H;
        P.e_b2 = P.e_w2 + H * H;
        P.e_w3 = P.e_b2 + H;
        P.e_b3 = P.e_w3 + 64 * H;
    }
    if (e == hipSuccess) e = song::launch_learned_forward(L, P, which, dX, n, dY, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(logits_out, dY, (size_t)n * out_dim * 4, hipMemcpyDeviceToHost);
    for (void* q : {(void*)dX, (void*)dW, (void*)dY})
        if (q) (void)hipFree(q);
    if (e != hipSuccess) return hip_fail(e, "mlp forward");
    return SONG_OK;
}

int song_index_bruteforce(song_index* ix, const void* queries, int32_t nq, int32_t k, int32_t* ids_out,
                          float* dists_out, double* kernel_ms_out) {
    if (!ix) return fail(SONG_ERR_ARG, "null index");
    if (k < 1 || k > 128) return fail(SONG_ERR_ARG, "k must be in [1, 128]");
    if (nq < 0 || (nq > 0 && (!queries || !ids_out))) return fail(SONG_ERR_ARG, "bad query buffers");
    if (nq == 0) return SONG_OK;
    if (ix->metric != SONG_METRIC_HAMMING && !rows_finite((const float*)queries, nq, ix->dim, ix->dim))
        return fail(SONG_ERR_ARG, "queries must be finite (NaN or Inf found)");
    std::lock_guard<std::mutex> lock(ix->mu);
    HIP_TRY(hipSetDevice(ix->device), "hipSetDevice");
    const size_t rb = row_bytes(ix);
    int qw = ix->metric == SONG_METRIC_HAMMING ? 8 : ix->dstride <= 256 ? 8 : ix->dstride <= 1024 ? 4 : 2;
    const int64_t groups = (nq + qw - 1) / qw;
    int64_t parts = (16384 + groups - 1) / groups;
    parts = std::max<int64_t>(1, std::min<int64_t>(parts, std::max<int64_t>(1, ix->n / 4096)));
    parts = std::min<int64_t>(parts, 65535);
    const int64_t slice = (ix->n + parts - 1) / parts;
    void* dq = nullptr;
    uint64_t* partial = nullptr;
    int32_t* dids = nullptr;
    float* dd = nullptr;
    auto cleanup = [&]() {
        if (dq) (void)hipFree(dq);
        if (partial) (void)hipFree(partial);
        if (dids) (void)hipFree(dids);
        if (dd) (void)hipFree(dd);
    };
    hipError_t e = hipMalloc(&dq, (size_t)nq * rb);
    if (e == hipSuccess) e = hipMalloc(&partial, (size_t)nq * parts * k * 8);
    if (e == hipSuccess) e = hipMalloc(&dids, (size_t)nq * k * 4);
    if (e == hipSuccess) e = hipMalloc(&dd, (size_t)nq * k * 4);
    if (e != hipSuccess) {
        cleanup();
        return hip_fail(e, "hipMalloc (bruteforce workspace)");
    }
    const void* src = stage_rows(ix, queries, nq);
    e = hipMemcpyAsync(dq, src, (size_t)nq * rb, hipMemcpyHostToDevice, ix->stream);
    song::GtArgs a = {};
    a.data = ix->d_data;
    a.queries = dq;
    a.partial = partial;
    a.ids_out = dids;
    a.dists_out = dd;
    a.n = ix->n;
    a.slice = slice;
    a.dstride = ix->dstride;
    a.nq = nq;
    a.k = k;
    a.parts = (int32_t)parts;
    if (e == hipSuccess) e = hipEventRecord(ix->ev[1], ix->stream);
    if (e == hipSuccess) e = song::launch_bruteforce(a, ix->metric, qw, ix->stream);
    if (e == hipSuccess) e = hipEventRecord(ix->ev[2], ix->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ids_out, dids, (size_t)nq * k * 4, hipMemcpyDeviceToHost, ix->stream);
    if (e == hipSuccess && dists_out)
        e = hipMemcpyAsync(dists_out, dd, (size_t)nq * k * 4, hipMemcpyDeviceToHost, ix->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ix->stream);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ix->ev[1], ix->ev[2]);
    cleanup();
    if (e != hipSuccess) return hip_fail(e, "bruteforce");
    if (kernel_ms_out) *kernel_ms_out = ms;
    return SONG_OK;
}

int song_bithash_project_device(const float* d_R, int32_t dim, const float* d_vecs, int64_t n, int32_t device,
                                uint64_t* d_codes_out, double* kernel_ms_out) {
    if (dim < 1 || dim > 2048) return fail(SONG_ERR_ARG, "bithash on the device supports dim in [1, 2048]");
    if (n < 0) return fail(SONG_ERR_ARG, "n must be >= 0");
    if (!d_R || (n > 0 && (!d_vecs || !d_codes_out))) return fail(SONG_ERR_ARG, "null buffer");
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a), "hipEventCreate");
    hipError_t e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        return hip_fail(e, "hipEventCreate");
    }
    e = hipEventRecord(a, 0);
    if (e == hipSuccess) e = song::launch_bithash(d_R, dim, d_vecs, n, d_codes_out, 0);
    if (e == hipSuccess) e = hipEventRecord(b, 0);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return hip_fail(e, "bithash projection");
    if (kernel_ms_out) *kernel_ms_out = ms;
    return SONG_OK;
}

int song_dev_alloc(int32_t device, int64_t bytes, void** out) {
    if (!out || bytes < 0) return fail(SONG_ERR_ARG, "bad dev_alloc arguments");
    *out = nullptr;
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    HIP_TRY(hipMalloc(out, bytes > 0 ? bytes : 1), "hipMalloc");
    return SONG_OK;
}

int song_dev_free(void* p) {
    if (p) HIP_TRY(hipFree(p), "hipFree");
    return SONG_OK;
}

int song_dev_copy(void* dst, const void* src, int64_t bytes, double* kernel_ms_out) {
    if (!dst || !src || bytes < 0 || (bytes % 16) != 0) return fail(SONG_ERR_ARG, "bytes must be a multiple of 16");
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a), "hipEventCreate");
    HIP_TRY(hipEventCreate(&b), "hipEventCreate");
    const int64_t n16 = bytes / 16;
    const int blocks = (int)std::min<int64_t>(8192, std::max<int64_t>(1, (n16 + 255) / 256));
    HIP_TRY(hipEventRecord(a, 0), "hipEventRecord");
    hipLaunchKernelGGL(copy16_kernel, dim3(blocks), dim3(256), 0, 0, (const uint4*)src, (uint4*)dst, n16);
    HIP_TRY(hipGetLastError(), "copy kernel launch");
    HIP_TRY(hipEventRecord(b, 0), "hipEventRecord");
    HIP_TRY(hipEventSynchronize(b), "hipEventSynchronize");
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (kernel_ms_out) *kernel_ms_out = ms;
    return SONG_OK;
}

int song_dev_gather(void* dst, const void* src, int64_t src_rows, int32_t row_bytes, const int32_t* ids,
                    int64_t n_ids, double* kernel_ms_out) {
    if (!dst || !src || !ids || n_ids < 1 || src_rows < 1) return fail(SONG_ERR_ARG, "bad gather arguments");
    if (row_bytes < 16 || row_bytes % 16 != 0) return fail(SONG_ERR_ARG, "row_bytes must be a positive multiple of 16");
    for (int64_t i = 0; i < n_ids; i++)
        if (ids[i] < 0 || ids[i] >= src_rows) return fail(SONG_ERR_ARG, "gather id out of range");
    int32_t* d_ids = nullptr;
    HIP_TRY(hipMalloc(&d_ids, (size_t)n_ids * 4), "hipMalloc gather ids");
    hipEvent_t a = nullptr, b = nullptr;
    hipError_t e = hipMemcpy(d_ids, ids, (size_t)n_ids * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&a);
    if (e == hipSuccess) e = hipEventCreate(&b);
    const int chunks = row_bytes / 16;
    const int blocks = (int)std::min<int64_t>(8192, std::max<int64_t>(1, (n_ids * chunks + 255) / 256));
    if (e == hipSuccess) e = hipEventRecord(a, 0);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(gather16_kernel, dim3(blocks), dim3(256), 0, 0, (const uint4*)src, d_ids, (uint4*)dst,
                           n_ids, chunks);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(b, 0);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(d_ids);
    if (e != hipSuccess) return hip_fail(e, "gather kernel");
    if (kernel_ms_out) *kernel_ms_out = ms;
    return SONG_OK;
}

int song_memcpy_h2d(void* dst, const void* src, int64_t bytes) {
    if ((!dst || !src) && bytes > 0) return fail(SONG_ERR_ARG, "null buffer");
    if (bytes > 0) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    return SONG_OK;
}

int song_memcpy_d2h(void* dst, const void* src, int64_t bytes) {
    if ((!dst || !src) && bytes > 0) return fail(SONG_ERR_ARG, "null buffer");
    if (bytes > 0) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return SONG_OK;
}

int song_queries_to_device(song_index* ix, const void* queries, int32_t nq, void** d_out) {
    if (!ix || !d_out || (nq > 0 && !queries) || nq < 0) return fail(SONG_ERR_ARG, "bad arguments");
    if (ix->metric != SONG_METRIC_HAMMING && !rows_finite((const float*)queries, nq, ix->dim, ix->dim))
        return fail(SONG_ERR_ARG, "queries must be finite (NaN or Inf found)");
    std::lock_guard<std::mutex> lock(ix->mu);
    HIP_TRY(hipSetDevice(ix->device), "hipSetDevice");
    const size_t rb = row_bytes(ix);
    *d_out = nullptr;
    HIP_TRY(hipMalloc(d_out, std::max<size_t>(1, (size_t)nq * rb)), "hipMalloc queries");
    const void* src = stage_rows(ix, queries, nq);
    hipError_t e = hipMemcpy(*d_out, src, (size_t)nq * rb, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(*d_out);
        *d_out = nullptr;
        return hip_fail(e, "queries H2D");
    }
    return SONG_OK;
}

}  // extern "C"
