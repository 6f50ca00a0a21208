// song_plan.hip -- the planning pass of the learned search (no counterpart in
// the reference, which seeds every query at ids p * n / N_MULTIPROBE,
// kernalGraph.h:215-231, and runs queries in input order).
//
// One wave per query computes the 16-d query projection once and feeds it to
//  * the entry network: the n_entry best-scoring pivot nodes become the
//    query's seed list, which the search kernel reads like any caller-given
//    entry list;
//  * the cost network: a prediction of the query's (standardised) log2 hop
//    count.
// A single-workgroup counting sort then orders the queue by predicted cost,
// most expensive first, so the longest searches start first and overlap the
// short ones instead of trailing after them.  Queries are independent: the
// order changes when each one runs, never its result.
#include "song_search_impl.h"

namespace song {

namespace {

// One dense 64 -> 64 layer, one output unit per lane: lane j returns
// b[j] + sum_i W[j][i] x_i, where x_i is held by lane i.  W row-major in global
// memory (read-only, shared by every wave, so it stays in cache).
__device__ __forceinline__ float dense64_global(const float* __restrict__ W, const float* __restrict__ b, float x,
                                                int lane) {
    const float4* row = (const float4*)(W + lane * 64);
    float acc = b[lane];
#pragma unroll
    for (int i4 = 0; i4 < 16; i4++) {
        const float4 w = row[i4];
        acc = __builtin_fmaf(w.x, readlane_f(x, 4 * i4 + 0), acc);
        acc = __builtin_fmaf(w.y, readlane_f(x, 4 * i4 + 1), acc);
        acc = __builtin_fmaf(w.z, readlane_f(x, 4 * i4 + 2), acc);
        acc = __builtin_fmaf(w.w, readlane_f(x, 4 * i4 + 3), acc);
    }
    return acc;
}

// First layer on wave-uniform inputs z: lane j returns relu(b1[j] + W1[j] . z).
__device__ __forceinline__ float dense16_relu(const float* __restrict__ W1, const float* __restrict__ b1,
                                              const float (&z)[LM_P], int lane) {
    const float4* w1 = (const float4*)(W1 + lane * LM_P);
    float h = b1[lane];
#pragma unroll
    for (int p4 = 0; p4 < LM_P / 4; p4++) {
        const float4 w = w1[p4];
        h = __builtin_fmaf(w.x, z[4 * p4 + 0], h);
        h = __builtin_fmaf(w.y, z[4 * p4 + 1], h);
        h = __builtin_fmaf(w.z, z[4 * p4 + 2], h);
        h = __builtin_fmaf(w.w, z[4 * p4 + 3], h);
    }
    return fmaxf(h, 0.f);
}

// Entry network: lane c returns the score of pivot c.
__device__ __forceinline__ float entry_scores(const PlanArgs& P, const float (&z)[LM_P], int lane) {
    const float h = dense16_relu(P.e_w1, P.e_b1, z, lane);
    const float h2 = fmaxf(dense64_global(P.e_w2, P.e_b2, h, lane), 0.f);
    return dense64_global(P.e_w3, P.e_b3, h2, lane);
}

// Cost network (flat w1 [64 x 16] b1 w2 [64 x 64] b2 w3 [64] b3): wave-uniform output.
__device__ __forceinline__ float cost_output(const float* __restrict__ W, const float (&z)[LM_P], int lane) {
    constexpr int B1 = LM_H * LM_P, W2 = B1 + LM_H, B2 = W2 + LM_H * LM_H, W3 = B2 + LM_H, B3 = W3 + LM_H;
    const float h = dense16_relu(W, W + B1, z, lane);
    const float h2 = fmaxf(dense64_global(W + W2, W + B2, h, lane), 0.f);
    return wave_sum(h2 * W[W3 + lane]) + W[B3];
}

// z = proj_scale * (proj (q - mu)); Hamming codes enter as +-1 vectors of 256.
template <int METRIC>
__device__ __forceinline__ void project_query(const PlanArgs& P, int64_t q, int lane, float (&z)[LM_P]) {
#pragma unroll
    for (int p = 0; p < LM_P; p++) {
        float part = 0.f;
        if constexpr (METRIC == METRIC_HAMMING) {
            // lane l owns bits 4l..4l+3 of the code
            const uint64_t w = ((const uint64_t*)P.queries)[q * 4 + (lane >> 4)];
            const int b0 = (lane & 15) * 4;
            const float* pr = P.proj + p * 256 + lane * 4;
#pragma unroll
            for (int b = 0; b < 4; b++) part = __builtin_fmaf(pr[b], ((w >> (b0 + b)) & 1ull) ? 1.f : -1.f, part);
        } else {
            const float* x = (const float*)P.queries + q * P.dstride;
            const float* pr = P.proj + (size_t)p * P.dstride;
            for (int e = lane; e < P.dstride; e += 64) part = __builtin_fmaf(pr[e], x[e] - P.mu[e], part);
        }
        z[p] = wave_sum(part) * P.proj_scale[p];
    }
}

template <int METRIC>
__global__ void __launch_bounds__(256) plan_kernel(PlanArgs P) {
    const int lane = lane_id();
    const int64_t q = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= P.nq) return;
    float z[LM_P];
    project_query<METRIC>(P, q, lane, z);
    if (P.entries) {
        float score = entry_scores(P, z, lane);
        int sid = 0;
        // the row: n_ref reference seeds, then the best-scoring pivots
        const int nr = P.n_ref;
        if (__ballot(!(score == score))) {
            // a score that is not a number: the reference's evenly spaced seeds
            // (repeated to fill the row; repeats are visited-set hits)
            const int m = nr > 0 ? nr : P.n_entry;
            sid = (int)(((uint64_t)(lane % m) * (uint64_t)P.n) / (uint64_t)m);
        } else {
            if (lane < nr) sid = (int)(((uint64_t)lane * (uint64_t)P.n) / (uint64_t)nr);
            const int piv = P.pivots[lane];
            for (int r = nr; r < P.n_entry; r++) {
                const float m = wave_max(score);
                const int l = ffs64(__ballot(score == m));
                const int id = __builtin_amdgcn_readlane(piv, l);
                if (lane == r) sid = id;
                if (lane == l) score = -INFINITY;
            }
        }
        if (lane < P.n_entry) P.entries[q * P.n_entry + lane] = sid;
    }
    if (P.cost) {
        const float c = cost_output(P.c_weights, z, lane);
        if (lane == 0) P.cost[q] = c == c ? c : 0.f;
    }
}

constexpr int COST_BINS = 256;

// Bin of a predicted cost, most expensive first: [-4, 4) in 256 steps.
__device__ __forceinline__ int cost_bin(float c) {
    const float t = (c + 4.f) * (COST_BINS / 8.f);
    const int b = t >= (float)(COST_BINS - 1) ? COST_BINS - 1 : (t > 0.f ? (int)t : 0);
    return COST_BINS - 1 - b;
}

// Counting sort of nq queries by cost bin (one workgroup).  Within a bin the
// order follows the LDS atomics, which only decides when a query starts.
__global__ void __launch_bounds__(1024) order_by_cost_kernel(const float* __restrict__ cost, int nq,
                                                             int32_t* __restrict__ order) {
    __shared__ int hist[COST_BINS];
    for (int i = threadIdx.x; i < COST_BINS; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < nq; q += blockDim.x) atomicAdd(&hist[cost_bin(cost[q])], 1);
    __syncthreads();
    if (threadIdx.x < 64) {
        // exclusive scan over the 256 bins: 4 per lane, then across lanes
        const int l = threadIdx.x;
        int v[4], s = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            v[j] = s;
            s += hist[4 * l + j];
        }
        int incl = s;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int o = __shfl_up(incl, d, 64);
            if (l >= d) incl += o;
        }
        const int excl = incl - s;
#pragma unroll
        for (int j = 0; j < 4; j++) hist[4 * l + j] = excl + v[j];
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nq; q += blockDim.x) {
        const int pos = atomicAdd(&hist[cost_bin(cost[q])], 1);
        order[pos] = q;
    }
}

// The kernels' inference code on given inputs (checked against a plain FP32 MLP).
__global__ void __launch_bounds__(64) stop_forward_kernel(LearnedArgs L, const float* __restrict__ X, int64_t n,
                                                          float* __restrict__ Y) {
    __shared__ float lds[LM_STOP_FLOATS];
    __shared__ float hbuf[64];
    stage_stop_weights(L.s_weights, lds);
    __syncthreads();
    const int lane = lane_id();
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        float f[LM_F];
#pragma unroll
        for (int j = 0; j < LM_F; j++) f[j] = X[i * LM_F + j];
        const float y = stop_logit(lds, f, L.s_b3, lane, hbuf);
        if (lane == 0) Y[i] = y;
    }
}

__global__ void __launch_bounds__(64) query_net_forward_kernel(PlanArgs P, int which, const float* __restrict__ X,
                                                               int64_t n, float* __restrict__ Y) {
    const int lane = lane_id();
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        float z[LM_P];
#pragma unroll
        for (int j = 0; j < LM_P; j++) z[j] = X[i * LM_P + j];
        if (which == 1) {
            Y[i * LM_C + lane] = entry_scores(P, z, lane);
        } else {
            const float c = cost_output(P.c_weights, z, lane);
            if (lane == 0) Y[i] = c;
        }
    }
}

}  // namespace

hipError_t launch_plan(const PlanArgs& P, int metric, int32_t* order, hipStream_t st) {
    if (P.nq <= 0) return hipSuccess;
    if (order && !P.cost) return hipErrorInvalidValue;
    const int blocks = (P.nq + 3) / 4;
    if (metric == METRIC_HAMMING) hipLaunchKernelGGL(plan_kernel<METRIC_HAMMING>, dim3(blocks), dim3(256), 0, st, P);
    else hipLaunchKernelGGL(plan_kernel<METRIC_L2>, dim3(blocks), dim3(256), 0, st, P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !order) return e;
    hipLaunchKernelGGL(order_by_cost_kernel, dim3(1), dim3(1024), 0, st, P.cost, P.nq, order);
    return hipGetLastError();
}

hipError_t launch_learned_forward(const LearnedArgs& L, const PlanArgs& P, int which, const float* X, int64_t n,
                                  float* Y, hipStream_t st) {
    const int blocks = (int)(n < 4096 ? (n > 0 ? n : 1) : 4096);
    if (which == 0) hipLaunchKernelGGL(stop_forward_kernel, dim3(blocks), dim3(64), 0, st, L, X, n, Y);
    else hipLaunchKernelGGL(query_net_forward_kernel, dim3(blocks), dim3(64), 0, st, P, which, X, n, Y);
    return hipGetLastError();
}

}  // namespace song
