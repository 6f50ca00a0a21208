/*
 * song_oracle.c -- CPU restatement of the reference's SONG graph search.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP search
 * kernels are checked against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as a checker or as the
 * timed CPU baseline -- never as part of the product path.  The product
 * (libsong.so) does not link or call anything in here.
 *
 * What it restates (reference files, numbered per file):
 *   - Bloom visited set        bloomFilter.h:14-49
 *   - Candidates / Results     gpuHeap.h:8-109  (unsorted arrays, first-index
 *                              min/max scans, swap-with-last removal)
 *   - kg_search_kernel         kernalGraph.h:184-299 (control flow, seeds,
 *                              early stop, trim + drain, QueryTrace counters)
 *   - hkg_search_kernel        hashKernelGraph.h:238-352 (same, Hamming)
 *   - warp_l2 / warp_hamming   kernalGraph.h:157-170, hashKernelGraph.h:215-224
 *                              (values; the summation order is the one the
 *                              MI355X kernel uses, see or_row_dist below)
 *   - l2 (sequential sum)      distance.h:5-12
 *   - Graph::search            graph.h:33-115 (CPU baseline)
 *   - the MI355X search's learned stop (no reference counterpart): its
 *     per-hop features, consult schedule and network, as the kernel computes
 *     them (or_search_learned, or_learned_trace, or_stop_logits)
 *   - brute-force ground truth bench_cpu.cpp:158-175, run_compare.py:199-213
 *
 * Everything is parameterised where the reference has compile-time
 * constants (config.h:9-58): row width ("block"), Bloom size, seed count,
 * queue size (ef) and k.
 *
 * Pins (what checks this file against something it did not write itself):
 *   see tests/test_oracle_pins.py -- knob reduction to the reference's own
 *   Graph::search compiled from its sources, an independent Python
 *   restatement of the Bloom hash plus its false-positive rate against the
 *   closed form, a brute-force list model of the two sets, and brute-force
 *   nearest neighbours.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 * -ffp-contract=off matters: every fused multiply-add below is an explicit
 * fmaf(), every other product is rounded on its own, as on the GPU.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_METRIC_L2 0
#define OR_METRIC_IP 1
#define OR_METRIC_HAMMING 2

/* Trace field order of QueryTrace, telemetry.h:5-18. */
enum {
    TR_SEED = 0, TR_VISITED, TR_EXPANDED, TR_SLOTS, TR_CANDIDATE, TR_DEVALS,
    TR_FPUSH, TR_RUPD, TR_PEAKF, TR_PEAKR, TR_RCOUNT, TR_EARLY, TR_N
};

/* ------------------------------------------------------------------------- */
/* Bloom filter, bloomFilter.h:14-49                                          */
/* ------------------------------------------------------------------------- */

/* BF_RAND, bloomFilter.h:14-19. Hash h uses (R[2h], R[2h+1]). */
static const uint64_t OR_BF_RAND[8] = {
    0x4bcb391f924ed183ULL, 0xa0ab69ccd854fc0aULL,
    0x91086b9cecf5e3b7ULL, 0xc68e01641bead407ULL,
    0x3a7b976128a30449ULL, 0x6d122efabfc4d99fULL,
    0xe6700ef8715030e2ULL, 0x80dd0c3bffcfb45bULL,
};

/* bf_hash, bloomFilter.h:23-30: xorshift-multiply, then mod the bit count. */
int or_bloom_pos(int h, int64_t id, int64_t nbits)
{
    uint64_t x = (uint64_t)id;
    x ^= x >> 33;
    x *= OR_BF_RAND[h * 2];
    x ^= x >> 33;
    x *= OR_BF_RAND[h * 2 + 1];
    x ^= x >> 33;
    return (int)(x % (uint64_t)nbits);
}

/* bf_set, bloomFilter.h:34-39: word = pos % W, bit = pos / W. */
static void bf_set(uint64_t *bloom, int64_t words, int id)
{
    for (int h = 0; h < 4; h++) {
        int pos = or_bloom_pos(h, id, words * 64);
        bloom[pos % words] |= (1ULL << (pos / words));
    }
}

/* bf_test, bloomFilter.h:43-49. */
static int bf_test(const uint64_t *bloom, int64_t words, int id)
{
    for (int h = 0; h < 4; h++) {
        int pos = or_bloom_pos(h, id, words * 64);
        if (!((bloom[pos % words] >> (pos / words)) & 1ULL)) return 0;
    }
    return 1;
}

/* Exposed for tests: positions of the 4 hashes for each id. */
void or_bloom_positions(const int32_t *ids, int n, int64_t nbits, int32_t *pos_out)
{
    for (int i = 0; i < n; i++)
        for (int h = 0; h < 4; h++) pos_out[i * 4 + h] = or_bloom_pos(h, ids[i], nbits);
}

/* Exposed for tests: insert ids[0..n_insert) then report bf_test for probes. */
void or_bloom_run(int64_t words, const int32_t *ins, int n_ins, const int32_t *probe, int n_probe,
                  int32_t *hit_out, uint64_t *bits_out)
{
    uint64_t *b = (uint64_t *)calloc((size_t)words, 8);
    for (int i = 0; i < n_ins; i++) bf_set(b, words, ins[i]);
    for (int i = 0; i < n_probe; i++) hit_out[i] = bf_test(b, words, probe[i]);
    if (bits_out) memcpy(bits_out, b, (size_t)words * 8);
    free(b);
}

/* ------------------------------------------------------------------------- */
/* Candidates / Results, gpuHeap.h:8-109                                      */
/* ------------------------------------------------------------------------- */

typedef struct {
    float *d;
    int32_t *id;
    int sz, cap;
} or_set;

/* init, gpuHeap.h:15-18 / 63-66 (cap clamped to >= 1; the array is sized by
 * the caller, so the MAX_SIZE clamp is the caller's allocation). */
static void set_init(or_set *s, float *d, int32_t *id, int cap)
{
    s->d = d;
    s->id = id;
    s->sz = 0;
    s->cap = cap < 1 ? 1 : cap;
}

/* Candidates::push, gpuHeap.h:23-36 and Results::try_push, gpuHeap.h:78-87
 * (identical rule): append while not full; when full, find the FIRST index
 * of the maximum and replace it only if d is strictly smaller. */
static int set_push(or_set *s, float d, int node)
{
    if (s->sz < s->cap) {
        s->d[s->sz] = d;
        s->id[s->sz] = node;
        s->sz++;
        return 1;
    }
    int fi = 0;
    for (int i = 1; i < s->cap; i++)
        if (s->d[i] > s->d[fi]) fi = i;
    if (d < s->d[fi]) {
        s->d[fi] = d;
        s->id[fi] = node;
        return 1;
    }
    return 0;
}

/* Candidates::pop_min, gpuHeap.h:39-49: first index of the minimum, the
 * hole is filled with the last element. */
static int set_pop_min(or_set *s, float *out_d)
{
    int mi = 0;
    for (int i = 1; i < s->sz; i++)
        if (s->d[i] < s->d[mi]) mi = i;
    *out_d = s->d[mi];
    int node = s->id[mi];
    s->d[mi] = s->d[s->sz - 1];
    s->id[mi] = s->id[s->sz - 1];
    s->sz--;
    return node;
}

/* Results::worst, gpuHeap.h:71-75. */
static float set_worst(const or_set *s)
{
    float m = s->d[0];
    for (int i = 1; i < s->sz; i++)
        if (s->d[i] > m) m = s->d[i];
    return m;
}

/* Results::pop_worst, gpuHeap.h:90-96. */
static void set_pop_worst(or_set *s)
{
    int wi = 0;
    for (int i = 1; i < s->sz; i++)
        if (s->d[i] > s->d[wi]) wi = i;
    s->d[wi] = s->d[s->sz - 1];
    s->id[wi] = s->id[s->sz - 1];
    s->sz--;
}

/* Results::drain_best, gpuHeap.h:99-108: repeated first-index minimum with
 * swap-with-last, best first. Also returns the distances (extension). */
static int set_drain_best(or_set *s, int32_t *out, float *out_d, int k)
{
    int count = s->sz < k ? s->sz : k;
    for (int i = 0; i < count; i++) {
        int mi = 0;
        for (int j = 1; j < s->sz; j++)
            if (s->d[j] < s->d[mi]) mi = j;
        out[i] = s->id[mi];
        if (out_d) out_d[i] = s->d[mi];
        s->d[mi] = s->d[s->sz - 1];
        s->id[mi] = s->id[s->sz - 1];
        s->sz--;
    }
    return count;
}

/*
 * Exposed for the brute-force model test: run an op sequence on one set.
 * op: 0 = push(v, id) (Candidates::push == Results::try_push),
 *     1 = pop_min, 2 = pop_worst, 3 = worst, 4 = drain_best(k = ids[i]).
 * ret_out[i]: push -> accepted flag, pop_min -> node id, worst -> 0.
 * val_out[i]: pop_min -> popped distance, worst -> worst value.
 * Final contents (positions included) are written to fin_d / fin_id,
 * drained ids to drain_out.  Returns the final size.
 */
int or_set_run(int cap, const int32_t *ops, const float *vals, const int32_t *ids, int nops,
               int32_t *ret_out, float *val_out, float *fin_d, int32_t *fin_id,
               int32_t *drain_out, float *drain_d)
{
    float *d = (float *)calloc((size_t)(cap < 1 ? 1 : cap), sizeof(float));
    int32_t *id = (int32_t *)calloc((size_t)(cap < 1 ? 1 : cap), sizeof(int32_t));
    or_set s;
    set_init(&s, d, id, cap);
    int ndrain = 0;
    for (int i = 0; i < nops; i++) {
        ret_out[i] = 0;
        val_out[i] = 0.0f;
        switch (ops[i]) {
        case 0: ret_out[i] = set_push(&s, vals[i], ids[i]); break;
        case 1:
            if (s.sz > 0) ret_out[i] = set_pop_min(&s, &val_out[i]);
            else ret_out[i] = -1;
            break;
        case 2:
            if (s.sz > 0) set_pop_worst(&s);
            break;
        case 3:
            if (s.sz > 0) val_out[i] = set_worst(&s);
            break;
        case 4:
            ndrain = set_drain_best(&s, drain_out, drain_d, ids[i]);
            ret_out[i] = ndrain;
            break;
        }
    }
    memcpy(fin_d, d, (size_t)s.sz * sizeof(float));
    memcpy(fin_id, id, (size_t)s.sz * sizeof(int32_t));
    int sz = s.sz;
    free(d);
    free(id);
    return sz;
}

/* ------------------------------------------------------------------------- */
/* Distances                                                                  */
/* ------------------------------------------------------------------------- */

/* l2, distance.h:5-12: sequential float sum, product rounded separately. */
float or_l2_sequential(const float *a, const float *b, int dim)
{
    float sum = 0.0f;
    for (int i = 0; i < dim; i++) {
        float d = a[i] - b[i];
        float p = d * d;
        sum = sum + p;
    }
    return sum;
}

/*
 * The value warp_l2 (kernalGraph.h:157-170) computes, summed in the order the
 * MI355X kernel sums it.  Rows are read as 16-byte chunks of 4 floats;
 * dim is padded with zeros to a multiple of 4.  A group of G lanes (G a
 * power of two, 1..64) owns one row; lane t of the group accumulates chunks
 * c = t, t+G, t+2G, ... in increasing order, each chunk's 4 elements in
 * order, with acc = fmaf(diff, diff, acc) (L2) or acc = fmaf(q, x, acc) (IP).
 * The G partial sums are then combined by an xor butterfly: for s = 1, 2,
 * 4, ... < G every lane t takes acc[t] + acc[t ^ s].  Lane 0's value is the
 * distance; IP returns -<q, x> so that smaller is better everywhere.
 */
float or_row_dist_f32(int metric, const float *q, const float *x, int dim, int G)
{
    float acc[64];
    float nxt[64];
    int d4 = (dim + 3) / 4;
    for (int t = 0; t < G; t++) {
        float a = 0.0f;
        for (int c = t; c < d4; c += G) {
            for (int e = 0; e < 4; e++) {
                int i = 4 * c + e;
                float qv = i < dim ? q[i] : 0.0f;
                float xv = i < dim ? x[i] : 0.0f;
                if (metric == OR_METRIC_IP) {
                    a = fmaf(qv, xv, a);
                } else {
                    float df = qv - xv;
                    a = fmaf(df, df, a);
                }
            }
        }
        acc[t] = a;
    }
    for (int s = 1; s < G; s <<= 1) {
        for (int t = 0; t < G; t++) nxt[t] = acc[t] + acc[t ^ s];
        memcpy(acc, nxt, (size_t)G * sizeof(float));
    }
    return metric == OR_METRIC_IP ? -acc[0] : acc[0];
}

/* warp_hamming, hashKernelGraph.h:215-224, on 256-bit codes held as 4 x u64
 * (same bits as the reference's 8 x u32, little-endian). */
int or_hamming256(const uint64_t *a, const uint64_t *b)
{
    int d = 0;
    for (int w = 0; w < 4; w++) d += __builtin_popcountll(a[w] ^ b[w]);
    return d;
}

typedef struct {
    int metric;
    int dim;
    int G;
    int ref_order; /* 1: distance.h sequential sum (knob for the CPU pin) */
    const float *fdata;
    const uint64_t *cdata;
} or_space;

static float space_dist(const or_space *sp, const void *q, int64_t node)
{
    if (sp->metric == OR_METRIC_HAMMING)
        return (float)or_hamming256((const uint64_t *)q, sp->cdata + node * 4);
    const float *x = sp->fdata + node * (int64_t)sp->dim;
    if (sp->ref_order) {
        if (sp->metric == OR_METRIC_IP) {
            float s = 0.0f;
            for (int i = 0; i < sp->dim; i++) {
                float p = ((const float *)q)[i] * x[i];
                s = s + p;
            }
            return -s;
        }
        return or_l2_sequential((const float *)q, x, sp->dim);
    }
    return or_row_dist_f32(sp->metric, (const float *)q, x, sp->dim, sp->G);
}

/* ------------------------------------------------------------------------- */
/* The GPU search algorithm, kernalGraph.h:184-299 / hashKernelGraph.h:238-352 */
/* ------------------------------------------------------------------------- */

typedef struct {
    const int32_t *adj;
    int64_t n;
    int block;
    int k, ef;
    int64_t bloom_words;
    int n_seeds;
    int exact_visited; /* knob: exact visited set instead of the Bloom filter */
    int unbounded;     /* knob: frontier never evicts (cap = n + n_seeds)     */
    int max_hops;      /* learned search's hop cap (0 = none; no reference
                          counterpart): stop once expanded > max_hops        */
    int res_cap;       /* result-set capacity: ef with the reference's early
                          stop; below ef (learned search's result_cap, no
                          reference counterpart) without that test           */
    const struct or_stop *stop; /* learned stop network, or NULL            */
} or_params;

/* ------------------------------------------------------------------------- */
/* The learned stop of the MI355X search (no reference counterpart)           */
/* ------------------------------------------------------------------------- */

/* The stop network and its consult schedule, as the search kernel applies
 * them (song_search_impl.h, search_query's stop test and stop_logit):
 * w: [64 x 8] w1 | [64] b1 | [64 x 64] w2 (row j = hidden unit j) | [64] b2 |
 * [64] w3, then b3; features are consulted at hops min_hops (or interval when
 * min_hops is 0), then every interval hops -- every late_interval hops once a
 * consult falls at or after late_hops -- provided k distances are known. */
typedef struct or_stop {
    const float *w;
    float b3, dist_scale, threshold;
    int min_hops, interval, confirm;
    int late_hops, late_interval; /* from hop late_hops on, consult every late_interval (0: never) */
    /* recording instead of stopping (the kernel's trace mode): per consult
     * the 8 features and the k-th best distance into rec [max_evals x 9] */
    float *rec;
    int max_evals, n_rec;
    int entry_late_hops; /* > 0: the search starts at the reference seeds and the
                            entry list joins at this hop (the kernel's late entry) */
} or_stop;

/* The k smallest distances seen so far, ascending (k <= 64): a new distance
 * d < the k-th smallest goes after the entries <= d; `inserts` counts them. */
typedef struct {
    float v[64];
    int n, k, inserts;
} or_topk;

static float topk_dk(const or_topk *t) { return t->n >= t->k ? t->v[t->k - 1] : INFINITY; }

static int topk_insert(or_topk *t, float d)
{
    if (!(d < topk_dk(t))) return 0;
    int pos = 0;
    while (pos < t->n && t->v[pos] <= d) pos++;
    int last = t->n < t->k ? t->n : t->k - 1;
    for (int i = last; i > pos; i--) t->v[i] = t->v[i - 1];
    t->v[pos] = d;
    if (t->n < t->k) t->n++;
    t->inserts++;
    return 1;
}

/* The network's logit in the kernel's operation order: each hidden unit an
 * fmaf chain over its inputs in index order from its bias, ReLU, the output
 * products w3[j] * h2[j] rounded on their own and summed as a pairwise tree
 * over units (0+1, 2+3, ..., then pairs of pairs), plus b3. */
static float stop_logit(const or_stop *st, const float f[8])
{
    const float *w1 = st->w, *b1 = w1 + 64 * 8, *w2 = b1 + 64, *b2 = w2 + 64 * 64, *w3 = b2 + 64;
    float h1[64], t[64];
    for (int j = 0; j < 64; j++) {
        float h = b1[j];
        for (int i = 0; i < 8; i++) h = fmaf(w1[j * 8 + i], f[i], h);
        h1[j] = fmaxf(h, 0.f);
    }
    for (int j = 0; j < 64; j++) {
        float a = b2[j];
        for (int i = 0; i < 64; i++) a = fmaf(w2[j * 64 + i], h1[i], a);
        a = fmaxf(a, 0.f);
        t[j] = a * w3[j];
    }
    for (int w = 64; w > 1; w /= 2)
        for (int j = 0; j < w / 2; j++) t[j] = t[2 * j] + t[2 * j + 1];
    return t[0] + st->b3;
}

static float clamp8(float x) { return fminf(fmaxf(x, -8.f), 8.f); }

/* Exposed for tests: the stop network's logit for n feature rows [n x 8]. */
void or_stop_logits(const float *stop_w, float stop_b3, const float *features, int n, float *out)
{
    or_stop st = {stop_w, stop_b3, 1.f, 0.f, 0, 1, 1, 0, 1, NULL, 0, 0};
    for (int i = 0; i < n; i++) out[i] = stop_logit(&st, features + (int64_t)i * 8);
}

/* Seeds at the entry of the search, kernalGraph.h:215-231: n_seeds evenly
 * spaced ids (size_t)p * n / n_seeds (N_MULTIPROBE = 4 in config.h:34). */
int64_t or_seed_id(int p, int64_t n, int n_seeds)
{
    return (int64_t)(((uint64_t)p * (uint64_t)n) / (uint64_t)n_seeds);
}

typedef struct {
    uint64_t *bloom;
    uint8_t *vis;
    or_set cand, res;
    float *cd, *rd;
    int32_t *ci, *ri;
    or_topk topk;        /* learned stop: the k smallest distances seen  */
    int last_insert_hop; /* learned stop: hop of the last top-k insertion */
} or_work;

static int visited_test_set(const or_params *p, or_work *w, int id)
{
    /* returns 1 if id counts as already visited; otherwise marks it. */
    if (p->exact_visited) {
        if (w->vis[id]) return 1;
        w->vis[id] = 1;
        return 0;
    }
    if (bf_test(w->bloom, p->bloom_words, id)) return 1;
    bf_set(w->bloom, p->bloom_words, id);
    return 0;
}

static void peak_update(int32_t *tr, const or_work *w)
{
    if (tr[TR_PEAKF] < w->cand.sz) tr[TR_PEAKF] = w->cand.sz;
    if (tr[TR_PEAKR] < w->res.sz) tr[TR_PEAKR] = w->res.sz;
}

/* One query; entry = list of seed ids (NULL: the reference's seeds). */
static void search_one(const or_space *sp, const or_params *p, or_work *w, const void *q,
                       const int32_t *entry, int n_entry, const int32_t *late, int n_late, int late_hop,
                       int32_t *ids_out, float *d_out, int32_t *tr)
{
    memset(tr, 0, TR_N * sizeof(int32_t));
    memset(&w->topk, 0, sizeof w->topk);
    w->topk.k = p->k;
    w->last_insert_hop = 0;
    int frontier_cap = p->unbounded ? (int)(p->n + n_entry + 1) : p->ef;
    set_init(&w->cand, w->cd, w->ci, frontier_cap);
    set_init(&w->res, w->rd, w->ri, p->res_cap);
    if (p->exact_visited) memset(w->vis, 0, (size_t)p->n);
    else memset(w->bloom, 0, (size_t)p->bloom_words * 8);

    /* seeds: kernalGraph.h:215-231 (test, then set, push, try_push) */
    for (int s = 0; s < n_entry; s++) {
        int ep = entry[s];
        float dp = space_dist(sp, q, ep);
        if (!visited_test_set(p, w, ep)) {
            int fa = set_push(&w->cand, dp, ep);
            int ra = set_push(&w->res, dp, ep);
            if (p->stop && topk_insert(&w->topk, dp)) w->last_insert_hop = 0;
            tr[TR_SEED]++;
            tr[TR_VISITED]++;
            tr[TR_DEVALS]++;
            tr[TR_FPUSH] += fa;
            tr[TR_RUPD] += ra;
            peak_update(tr, w);
        }
    }

    /* main loop: kernalGraph.h:233-289 */
    const or_stop *st = p->stop;
    int next_eval = st ? (st->min_hops > 0 ? st->min_hops : st->interval) : 0;
    int votes = 0;
    float dk_prev = INFINITY;
    for (;;) {
        if (w->cand.sz == 0) break;
        float c_min;
        int cn = set_pop_min(&w->cand, &c_min);
        tr[TR_EXPANDED]++;
        if (p->res_cap == p->ef && w->res.sz == w->res.cap && c_min > set_worst(&w->res)) {
            tr[TR_EARLY] = 1;
            break;
        }
        if (p->max_hops > 0 && tr[TR_EXPANDED] > p->max_hops) {
            tr[TR_EARLY] = 3;
            break;
        }
        if (st) {
            const int hop = tr[TR_EXPANDED];
            int due = 0;
            if (hop == next_eval) {
                due = 1;
                next_eval += st->late_hops > 0 && hop >= st->late_hops ? st->late_interval : st->interval;
            }
            const float dk = topk_dk(&w->topk);
            if (dk < INFINITY && due && hop >= st->min_hops) {
                /* reciprocals rounded once, then each feature one product (the kernel's order) */
                const float isc = 1.0f / st->dist_scale, d1 = w->topk.v[0];
                const float ik = 1.0f / (float)p->k, ief = 1.0f / (float)p->ef;
                int closer = 0; /* frontier entries (the popped one removed) closer than the k-th best */
                for (int i = 0; i < w->cand.sz; i++) closer += w->cand.d[i] < dk;
                float f[8];
                f[0] = (float)hop * (1.0f / 64.0f);
                f[1] = (c_min - dk) * isc;
                f[2] = (dk - d1) * isc;
                f[3] = (float)(hop - w->last_insert_hop) * (1.0f / 16.0f);
                f[4] = (float)closer * ik;
                f[5] = (float)w->topk.inserts * ik;
                f[6] = (float)w->cand.sz * ief;
                f[7] = dk_prev < INFINITY ? (dk_prev - dk) * isc : 0.f;
                for (int i = 0; i < 8; i++) f[i] = clamp8(f[i]);
                dk_prev = dk;
                w->topk.inserts = 0;
                if (st->rec) {
                    if (st->n_rec < st->max_evals) {
                        float *row = st->rec + (int64_t)st->n_rec * 9;
                        for (int i = 0; i < 8; i++) row[i] = f[i];
                        row[8] = dk;
                    }
                    ((or_stop *)st)->n_rec++;
                } else {
                    const float logit = stop_logit(st, f);
                    votes = logit > st->threshold ? votes + 1 : 0; /* a NaN logit never votes */
                    if (votes >= st->confirm) {
                        tr[TR_EARLY] = 2;
                        break;
                    }
                }
            }
        }
        if (late && tr[TR_EXPANDED] == late_hop) {
            /* the entry list joins: tested and pushed in order, before the popped node's row */
            for (int s = 0; s < n_late; s++) {
                int ep = late[s];
                if (ep < 0 || ep >= p->n) continue;
                if (visited_test_set(p, w, ep)) continue;
                float dp = space_dist(sp, q, ep);
                int fa = set_push(&w->cand, dp, ep);
                int ra = set_push(&w->res, dp, ep);
                if (st && topk_insert(&w->topk, dp)) w->last_insert_hop = tr[TR_EXPANDED];
                tr[TR_VISITED]++;
                tr[TR_CANDIDATE]++;
                tr[TR_DEVALS]++;
                tr[TR_FPUSH] += fa;
                tr[TR_RUPD] += ra;
                peak_update(tr, w);
            }
        }
        for (int j = 0; j < p->block; j++) {
            tr[TR_SLOTS]++;
            int nb = p->adj[(int64_t)cn * p->block + j];
            if (nb < 0) continue;
            if (visited_test_set(p, w, nb)) continue;
            tr[TR_VISITED]++;
            tr[TR_CANDIDATE]++;
            float nd = space_dist(sp, q, nb);
            tr[TR_DEVALS]++;
            int fa = set_push(&w->cand, nd, nb);
            int ra = set_push(&w->res, nd, nb);
            if (st && topk_insert(&w->topk, nd)) w->last_insert_hop = tr[TR_EXPANDED];
            tr[TR_FPUSH] += fa;
            tr[TR_RUPD] += ra;
            peak_update(tr, w);
        }
    }

    /* trim to k and drain best-first: kernalGraph.h:292-298 */
    while (w->res.sz > p->k) set_pop_worst(&w->res);
    int cnt = set_drain_best(&w->res, ids_out, d_out, p->k);
    for (int i = cnt; i < p->k; i++) {
        ids_out[i] = -1;
        if (d_out) d_out[i] = INFINITY;
    }
    tr[TR_RCOUNT] = cnt;
}

static int64_t row_elems(const or_space *sp) { return sp->metric == OR_METRIC_HAMMING ? 4 : sp->dim; }

/*
 * Full re-execution of the reference GPU search for nq queries.
 * data: float [n x dim] (L2/IP) or uint64 [n x 4] (Hamming).
 * queries: same element type, [nq x dim] or [nq x 4].
 * entries: optional [nq x n_entry] per-query seed lists (learned entry);
 *          NULL uses the reference seeds.
 * Returns 0, or -1 on bad arguments.
 */
static int search_all(const void *data, const int32_t *adj, int64_t n, int dim, int block, int metric,
                      const void *queries, int nq, int k, int ef, int64_t bloom_bits, int n_seeds, int G,
                      int ref_order, int exact_visited, int unbounded, const int32_t *entries, int n_entry,
                      int max_hops, int result_cap, const or_stop *stop, int32_t *ids_out, float *dists_out,
                      int32_t *trace_out);

int or_search_capped(const void *data, const int32_t *adj, int64_t n, int dim, int block, int metric,
                     const void *queries, int nq, int k, int ef, int64_t bloom_bits, int n_seeds, int G,
                     int ref_order, int exact_visited, int unbounded, const int32_t *entries, int n_entry,
                     int max_hops, int result_cap, int32_t *ids_out, float *dists_out, int32_t *trace_out);

int or_search(const void *data, const int32_t *adj, int64_t n, int dim, int block, int metric,
              const void *queries, int nq, int k, int ef, int64_t bloom_bits, int n_seeds, int G,
              int ref_order, int exact_visited, int unbounded, const int32_t *entries, int n_entry,
              int32_t *ids_out, float *dists_out, int32_t *trace_out)
{
    return or_search_capped(data, adj, n, dim, block, metric, queries, nq, k, ef, bloom_bits, n_seeds, G,
                            ref_order, exact_visited, unbounded, entries, n_entry, 0, 0, ids_out, dists_out,
                            trace_out);
}

/* or_search with the learned search's hop cap (max_hops > 0) and result-set
 * capacity (result_cap: 0 = ef; k <= result_cap < ef replaces the reference's
 * "best candidate worse than the worst result" test, which needs an ef-sized
 * result set, by the hop cap, which must then be set). */
int or_search_capped(const void *data, const int32_t *adj, int64_t n, int dim, int block, int metric,
                     const void *queries, int nq, int k, int ef, int64_t bloom_bits, int n_seeds, int G,
                     int ref_order, int exact_visited, int unbounded, const int32_t *entries, int n_entry,
                     int max_hops, int result_cap, int32_t *ids_out, float *dists_out, int32_t *trace_out)
{
    return search_all(data, adj, n, dim, block, metric, queries, nq, k, ef, bloom_bits, n_seeds, G, ref_order,
                      exact_visited, unbounded, entries, n_entry, max_hops, result_cap, NULL, ids_out, dists_out,
                      trace_out);
}

/* or_search_capped with the learned stop network (the MI355X search's learned
 * bit 1): stop_w is the network in the host layout (w1 [64 x 8], b1, w2
 * [64 x 64], b2, w3), consulted as or_stop describes; k <= 64. */
int or_search_learned(const void *data, const int32_t *adj, int64_t n, int dim, int block, int metric,
                      const void *queries, int nq, int k, int ef, int64_t bloom_bits, int n_seeds, int G,
                      const int32_t *entries, int n_entry, int max_hops, int result_cap, const float *stop_w,
                      float stop_b3, float dist_scale, float threshold, int min_hops, int interval, int confirm,
                      int late_hops, int late_interval, int entry_late_hops, int32_t *ids_out, float *dists_out,
                      int32_t *trace_out)
{
    if (!stop_w || k > 64 || interval < 1 || min_hops < 0 || confirm < 1 || !(dist_scale > 0.f)) return -1;
    if (late_hops < 0 || (late_hops > 0 && late_interval < 1) || entry_late_hops < 0) return -1;
    if (entry_late_hops > 0 && !entries) return -1;
    or_stop st = {stop_w, stop_b3, dist_scale, threshold, min_hops, interval, confirm, late_hops, late_interval,
                  NULL, 0, 0, entry_late_hops};
    return search_all(data, adj, n, dim, block, metric, queries, nq, k, ef, bloom_bits, n_seeds, G, 0, 0, 0,
                      entries, n_entry, max_hops, result_cap, &st, ids_out, dists_out, trace_out);
}

/* The features the learned search's trace mode records (song_learn_trace):
 * the search with the stop network never consulted for a decision, and at
 * each scheduled consult the 8 clamped features and the k-th best distance,
 * rec [nq x max_evals x 9]; rec_count [nq] = rows written; final_dk [nq] =
 * the k-th best distance at the end. */
int or_learned_trace(const void *data, const int32_t *adj, int64_t n, int dim, int block, int metric,
                     const void *queries, int nq, int k, int ef, int64_t bloom_bits, int n_seeds, int G,
                     const int32_t *entries, int n_entry, float dist_scale, int min_hops, int interval,
                     int late_hops, int late_interval, int max_evals, float *rec, int32_t *rec_count,
                     float *final_dk)
{
    if (k > 64 || interval < 1 || min_hops < 0 || max_evals < 1 || !(dist_scale > 0.f) || nq < 0) return -1;
    if (late_hops < 0 || (late_hops > 0 && late_interval < 1)) return -1;
    int32_t *ids = (int32_t *)malloc((size_t)k * sizeof(int32_t));
    float *d = (float *)malloc((size_t)k * sizeof(float));
    int32_t tr[TR_N];
    int rc = 0;
    const int64_t qstride = (metric == OR_METRIC_HAMMING ? 4 : dim) * (int64_t)(metric == OR_METRIC_HAMMING ? 8 : 4);
    for (int qi = 0; qi < nq && rc == 0; qi++) {
        or_stop st = {NULL, 0.f, dist_scale, 0.f, min_hops, interval, 1, late_hops, late_interval,
                      rec + (int64_t)qi * max_evals * 9, max_evals, 0};
        static const float none[4801];
        st.w = none;
        rc = search_all(data, adj, n, dim, block, metric, (const char *)queries + qi * qstride, 1, k, ef,
                        bloom_bits, n_seeds, G, 0, 0, 0, entries ? entries + (int64_t)qi * n_entry : NULL, n_entry,
                        0, 0, &st, ids, d, tr);
        rec_count[qi] = st.n_rec < max_evals ? st.n_rec : max_evals;
        final_dk[qi] = tr[TR_RCOUNT] >= k ? d[k - 1] : INFINITY;
    }
    free(ids);
    free(d);
    return rc;
}

static int search_all(const void *data, const int32_t *adj, int64_t n, int dim, int block, int metric,
                      const void *queries, int nq, int k, int ef, int64_t bloom_bits, int n_seeds, int G,
                      int ref_order, int exact_visited, int unbounded, const int32_t *entries, int n_entry,
                      int max_hops, int result_cap, const or_stop *stop, int32_t *ids_out, float *dists_out,
                      int32_t *trace_out)
{
    if (max_hops < 0 || result_cap < 0) return -1;
    if (n < 1 || k < 1 || ef < 1 || block < 1 || nq < 0 || G < 1 || G > 64) return -1;
    if (!exact_visited && (bloom_bits < 64 || (bloom_bits % 64) != 0)) return -1;
    if (ef < k) ef = k; /* kernalGraph.h:315 (the reference also caps ef at 50; lifted here) */
    int res_cap = ef;
    if (result_cap > 0 && result_cap < ef) {
        if (result_cap < k || max_hops == 0) return -1;
        res_cap = result_cap;
    }
    or_space sp = {metric, dim, G, ref_order, (const float *)data, (const uint64_t *)data};
    or_params p = {adj, n, block, k, ef, bloom_bits / 64, n_seeds, exact_visited, unbounded, max_hops, res_cap, stop};
    int ne = entries ? n_entry : n_seeds;
    int64_t fcap = unbounded ? n + ne + 1 : ef;
    or_work w;
    w.bloom = (uint64_t *)calloc((size_t)(bloom_bits / 64 + 1), 8);
    w.vis = (uint8_t *)calloc((size_t)n, 1);
    w.cd = (float *)malloc((size_t)fcap * sizeof(float));
    w.ci = (int32_t *)malloc((size_t)fcap * sizeof(int32_t));
    w.rd = (float *)malloc((size_t)ef * sizeof(float));
    w.ri = (int32_t *)malloc((size_t)ef * sizeof(int32_t));
    int32_t *seeds = (int32_t *)malloc((size_t)(ne > 0 ? ne : 1) * sizeof(int32_t));
    int64_t qstride = row_elems(&sp);
    int esz = metric == OR_METRIC_HAMMING ? 8 : 4;
    for (int qi = 0; qi < nq; qi++) {
        if (entries) {
            memcpy(seeds, entries + (int64_t)qi * ne, (size_t)ne * sizeof(int32_t));
        } else {
            for (int s = 0; s < ne; s++) seeds[s] = (int32_t)or_seed_id(s, n, ne);
        }
        const char *q = (const char *)queries + (int64_t)qi * qstride * esz;
        const int late_hop = stop && entries ? stop->entry_late_hops : 0;
        if (late_hop > 0) /* reference seeds first, the query's entry list at late_hop */
            for (int s = 0; s < ne; s++) seeds[s] = (int32_t)or_seed_id(s, n, ne);
        search_one(&sp, &p, &w, q, seeds, ne, late_hop > 0 ? entries + (int64_t)qi * ne : NULL, ne, late_hop,
                   ids_out + (int64_t)qi * k,
                   dists_out ? dists_out + (int64_t)qi * k : NULL, trace_out + (int64_t)qi * TR_N);
    }
    free(seeds);
    free(w.bloom);
    free(w.vis);
    free(w.cd);
    free(w.ci);
    free(w.rd);
    free(w.ri);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Graph::search, graph.h:33-115 -- the reference's CPU search (baseline)     */
/* ------------------------------------------------------------------------- */

/* Binary heap over (d, id) pairs, ordered lexicographically.  Graph::search
 * uses std::priority_queue<pair<float,int>> with the pair's total order and
 * ids are never pushed twice, so any correct heap pops the same sequence. */
typedef struct {
    float *d;
    int32_t *id;
    int64_t sz;
    int is_max;
} or_heap;

static int hp_less(const or_heap *h, int64_t a, int64_t b)
{
    /* "a before b" at the top */
    float da = h->d[a], db = h->d[b];
    int32_t ia = h->id[a], ib = h->id[b];
    if (h->is_max) return da > db || (da == db && ia > ib);
    return da < db || (da == db && ia < ib);
}

static void hp_swap(or_heap *h, int64_t a, int64_t b)
{
    float td = h->d[a];
    int32_t ti = h->id[a];
    h->d[a] = h->d[b];
    h->id[a] = h->id[b];
    h->d[b] = td;
    h->id[b] = ti;
}

static void hp_push(or_heap *h, float d, int32_t id)
{
    int64_t i = h->sz++;
    h->d[i] = d;
    h->id[i] = id;
    while (i > 0) {
        int64_t par = (i - 1) / 2;
        if (!hp_less(h, i, par)) break;
        hp_swap(h, i, par);
        i = par;
    }
}

static void hp_pop(or_heap *h)
{
    h->sz--;
    if (h->sz == 0) return;
    h->d[0] = h->d[h->sz];
    h->id[0] = h->id[h->sz];
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, best = i;
        if (l < h->sz && hp_less(h, l, best)) best = l;
        if (r < h->sz && hp_less(h, r, best)) best = r;
        if (best == i) break;
        hp_swap(h, i, best);
        i = best;
    }
}

/*
 * Graph::search (graph.h:33-115), L2 with the sequential distance.h sum.
 * Ids are written worst-first as Graph::search returns them, -1 padded.
 */
int or_cpu_search(const float *data, const int32_t *adj, int64_t n, int dim, int block,
                  const float *queries, int nq, int k, int ef, int n_seeds, int32_t *ids_out,
                  int32_t *trace_out)
{
    if (ef < k) ef = k; /* graph.h:34 */
    int64_t cap = n + 8;
    or_heap q = {(float *)malloc((size_t)cap * 4), (int32_t *)malloc((size_t)cap * 4), 0, 0};
    or_heap top = {(float *)malloc((size_t)(ef + 2) * 4), (int32_t *)malloc((size_t)(ef + 2) * 4), 0, 1};
    uint8_t *vis = (uint8_t *)calloc((size_t)n, 1);
    for (int qi = 0; qi < nq; qi++) {
        const float *qv = queries + (int64_t)qi * dim;
        int32_t *tr = trace_out + (int64_t)qi * TR_N;
        memset(tr, 0, TR_N * sizeof(int32_t));
        memset(vis, 0, (size_t)n);
        q.sz = 0;
        top.sz = 0;
        for (int p = 0; p < n_seeds; p++) { /* graph.h:48-62 */
            int64_t ep = or_seed_id(p, n, n_seeds);
            if (vis[ep]) continue;
            float d = or_l2_sequential(qv, data + ep * dim, dim);
            hp_push(&q, d, (int32_t)ep);
            hp_push(&top, d, (int32_t)ep);
            vis[ep] = 1;
            tr[TR_SEED]++;
            tr[TR_VISITED]++;
            tr[TR_DEVALS]++;
            tr[TR_FPUSH]++;
            tr[TR_RUPD]++;
            if (tr[TR_PEAKF] < q.sz) tr[TR_PEAKF] = (int32_t)q.sz;
            if (tr[TR_PEAKR] < top.sz) tr[TR_PEAKR] = (int32_t)top.sz;
        }
        while (q.sz > 0) { /* graph.h:64-106 */
            float cd = q.d[0];
            int32_t cu = q.id[0];
            hp_pop(&q);
            tr[TR_EXPANDED]++;
            if (top.sz == ef && cd > top.d[0]) {
                tr[TR_EARLY] = 1;
                break;
            }
            int32_t cands[64];
            int nc = 0;
            for (int j = 0; j < block; j++) {
                tr[TR_SLOTS]++;
                int32_t v = adj[(int64_t)cu * block + j];
                if (v < 0 || vis[v]) continue;
                vis[v] = 1;
                cands[nc++] = v;
                tr[TR_VISITED]++;
                tr[TR_CANDIDATE]++;
            }
            float dists[64];
            for (int i = 0; i < nc; i++) {
                dists[i] = or_l2_sequential(qv, data + (int64_t)cands[i] * dim, dim);
                tr[TR_DEVALS]++;
            }
            for (int i = 0; i < nc; i++) {
                hp_push(&q, dists[i], cands[i]);
                tr[TR_FPUSH]++;
                if (tr[TR_PEAKF] < q.sz) tr[TR_PEAKF] = (int32_t)q.sz;
                if (top.sz < ef || dists[i] < top.d[0]) {
                    hp_push(&top, dists[i], cands[i]);
                    if (top.sz > ef) hp_pop(&top);
                    tr[TR_RUPD]++;
                    if (tr[TR_PEAKR] < top.sz) tr[TR_PEAKR] = (int32_t)top.sz;
                }
            }
        }
        while (top.sz > k) hp_pop(&top); /* graph.h:108-114 */
        int32_t *out = ids_out + (int64_t)qi * k;
        int cnt = 0;
        while (top.sz > 0) {
            out[cnt++] = top.id[0];
            hp_pop(&top);
        }
        tr[TR_RCOUNT] = cnt;
        for (int i = cnt; i < k; i++) out[i] = -1;
    }
    free(q.d);
    free(q.id);
    free(top.d);
    free(top.id);
    free(vis);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Brute-force ground truth (bench_cpu.cpp:158-175; run_compare.py:199-213)   */
/* ------------------------------------------------------------------------- */

/* Exact k nearest by (distance, id) ascending, distances in double (L2/IP)
 * or integer popcount (Hamming).  Ids best-first. */
int or_bruteforce(const void *base, int64_t n, const void *queries, int nq, int dim, int metric,
                  int k, int32_t *ids_out, double *d_out)
{
    double *bd = (double *)malloc((size_t)k * sizeof(double));
    int32_t *bi = (int32_t *)malloc((size_t)k * sizeof(int32_t));
    for (int qi = 0; qi < nq; qi++) {
        int cnt = 0;
        for (int64_t j = 0; j < n; j++) {
            double d;
            if (metric == OR_METRIC_HAMMING) {
                d = or_hamming256((const uint64_t *)queries + (int64_t)qi * 4,
                                  (const uint64_t *)base + j * 4);
            } else {
                const float *q = (const float *)queries + (int64_t)qi * dim;
                const float *x = (const float *)base + j * dim;
                double s = 0.0;
                for (int e = 0; e < dim; e++) {
                    if (metric == OR_METRIC_IP) s += (double)q[e] * (double)x[e];
                    else {
                        double df = (double)q[e] - (double)x[e];
                        s += df * df;
                    }
                }
                d = metric == OR_METRIC_IP ? -s : s;
            }
            /* insertion into the sorted top-k (ties: smaller id first) */
            if (cnt == k && !(d < bd[k - 1] || (d == bd[k - 1] && j < bi[k - 1]))) continue;
            int pos = cnt < k ? cnt : k - 1;
            while (pos > 0 && (d < bd[pos - 1] || (d == bd[pos - 1] && j < bi[pos - 1]))) {
                bd[pos] = bd[pos - 1];
                bi[pos] = bi[pos - 1];
                pos--;
            }
            bd[pos] = d;
            bi[pos] = (int32_t)j;
            if (cnt < k) cnt++;
        }
        for (int i = 0; i < k; i++) {
            ids_out[(int64_t)qi * k + i] = i < cnt ? bi[i] : -1;
            if (d_out) d_out[(int64_t)qi * k + i] = i < cnt ? bd[i] : INFINITY;
        }
    }
    free(bd);
    free(bi);
    return 0;
}
