// song_gt.hip -- exact k-nearest-neighbour ground truth by brute force on the
// GPU, over an index's resident vectors.  Replaces the reference's CPU/numpy
// ground truth (bench_cpu.cpp:158-175, run_compare.py:199-213).
//
// Results are ordered by the 64-bit key (order-preserving distance bits, id),
// so ties between equal distances resolve to the smaller id and the answer is
// unique.  Distances are computed directly as sum((q - x)^2) in fp32 (exact for
// integer-valued vectors while the sum stays below 2^24), -<q, x> for IP, and
// popcount for Hamming.
//
// Pass 1 (gt_partial_kernel): one wave per (group of QW queries, slice of the
// base).  Queries sit in LDS; lane l computes the distances of row r0 + 64i + l
// to all QW queries, then each query's running top-k (keys in VGPRs, entry i
// in slot i/64 of lane i%64) admits the lanes that beat its current maximum, in
// lane order.  Pass 2 (gt_merge_kernel): one wave per query merges the slices'
// lists and writes them sorted.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "song_gt.h"

namespace song {

namespace {

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
    uint32_t lo = __shfl_xor((uint32_t)v, m, 64);
    uint32_t hi = __shfl_xor((uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        uint64_t o = shfl_xor_u64(v, m);
        v = o > v ? o : v;
    }
    return v;
}

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        uint64_t o = shfl_xor_u64(v, m);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ uint32_t order_bits(float d) {
    uint32_t b = __float_as_uint(d + 0.0f);  // -0 -> +0
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float unorder_bits(uint32_t o) {
    uint32_t b = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return __uint_as_float(b);
}

constexpr uint64_t KEY_NONE = ~0ull;

// Bounded set of the k smallest unique keys.
template <int S>
struct KeySet {
    typedef unsigned long long kvec __attribute__((ext_vector_type(S == 1 ? 2 : S)));
    kvec key;
    int sz, cap;
    uint64_t maxk;  // wave-uniform, valid when sz == cap

    __device__ __forceinline__ void init(int c) {
        sz = 0;
        cap = c;
        maxk = KEY_NONE;
#pragma unroll
        for (int s = 0; s < S; s++) key[s] = KEY_NONE;
    }
    __device__ __forceinline__ uint64_t lane_max(int lane) const {
        uint64_t m = 0;
#pragma unroll
        for (int s = 0; s < S; s++)
            if (s * 64 + lane < sz && key[s] > m) m = key[s];
        return m;
    }
    __device__ __forceinline__ void refresh_max(int lane) { maxk = wave_max_u64(lane_max(lane)); }
    // Admit the keys of lanes 0..63 (KEY_NONE = no candidate), in lane order:
    // append while not full, then replace the maximum when strictly smaller.
    __device__ __forceinline__ void push_lanes(uint64_t k, int lane) {
        while (sz < cap) {
            const uint64_t m = __ballot(k != KEY_NONE);
            if (m == 0ull) return;
            const int r = __ffsll((unsigned long long)m) - 1;
            const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)k, r);
            const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(k >> 32), r);
            const int slot = sz >> 6;
            if (lane == (sz & 63)) {
#pragma unroll
                for (int s = 0; s < S; s++)
                    if (s == slot) key[s] = ((uint64_t)hi << 32) | lo;
            }
            if (lane == r) k = KEY_NONE;
            sz++;
            if (sz == cap) refresh_max(lane);
        }
        for (;;) {
            const uint64_t m = __ballot(k < maxk);
            if (m == 0ull) break;
            const int r = __ffsll((unsigned long long)m) - 1;
            const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)k, r);
            const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(k >> 32), r);
            const uint64_t nk = ((uint64_t)hi << 32) | lo;
            // replace the entry holding maxk (keys are unique)
#pragma unroll
            for (int s = 0; s < S; s++)
                if (s * 64 + lane < sz && key[s] == maxk) key[s] = nk;
            if (lane == r) k = KEY_NONE;
            refresh_max(lane);
        }
    }
};

template <int METRIC, int S, int QW>
__global__ void __launch_bounds__(64) gt_partial_kernel(GtArgs a) {
    extern __shared__ float4 qs[];  // [QW][d4] query chunks (Hamming: 2 x ulonglong2 per query)
    const int lane = __lane_id();
    const int qg = blockIdx.x;      // query group
    const int part = blockIdx.y;    // base slice
    const int d4 = a.dstride / 4;
    const int q0 = qg * QW;
    const int nqg = min(QW, a.nq - q0);
    const int64_t r0 = (int64_t)part * a.slice;
    const int64_t r1 = min<int64_t>(a.n, r0 + a.slice);
    if constexpr (METRIC == METRIC_HAMMING) {
        const ulonglong2* src = (const ulonglong2*)a.queries;
        ulonglong2* dst = (ulonglong2*)qs;
        for (int i = lane; i < QW * 2; i += 64) dst[i] = (i / 2) < nqg ? src[(size_t)(q0 + i / 2) * 2 + (i & 1)] : make_ulonglong2(0, 0);
    } else {
        const float4* src = (const float4*)a.queries;
        for (int i = lane; i < QW * d4; i += 64) {
            const int qi = i / d4, c = i % d4;
            qs[i] = qi < nqg ? src[(size_t)(q0 + qi) * d4 + c] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
    __syncthreads();

    KeySet<S> sets[QW];
#pragma unroll
    for (int qi = 0; qi < QW; qi++) sets[qi].init(a.k);

    for (int64_t b = r0; b < r1; b += 64) {
        const int64_t r = b + lane;
        const bool valid = r < r1;
        float acc[QW];
#pragma unroll
        for (int qi = 0; qi < QW; qi++) acc[qi] = 0.f;
        if (valid) {
            if constexpr (METRIC == METRIC_HAMMING) {
                const ulonglong2* row = (const ulonglong2*)a.data + (size_t)r * 2;
                const ulonglong2 x0 = row[0], x1 = row[1];
                const ulonglong2* qq = (const ulonglong2*)qs;
#pragma unroll
                for (int qi = 0; qi < QW; qi++) {
                    const ulonglong2 y0 = qq[qi * 2], y1 = qq[qi * 2 + 1];
                    acc[qi] = (float)(__popcll(x0.x ^ y0.x) + __popcll(x0.y ^ y0.y) + __popcll(x1.x ^ y1.x) +
                                      __popcll(x1.y ^ y1.y));
                }
            } else {
                const float4* row = (const float4*)a.data + (size_t)r * d4;
                for (int c = 0; c < d4; c++) {
                    const float4 x = row[c];
#pragma unroll
                    for (int qi = 0; qi < QW; qi++) {
                        const float4 q = qs[qi * d4 + c];
                        if constexpr (METRIC == METRIC_IP) {
                            acc[qi] = __builtin_fmaf(q.x, x.x, acc[qi]);
                            acc[qi] = __builtin_fmaf(q.y, x.y, acc[qi]);
                            acc[qi] = __builtin_fmaf(q.z, x.z, acc[qi]);
                            acc[qi] = __builtin_fmaf(q.w, x.w, acc[qi]);
                        } else {
                            float e;
                            e = q.x - x.x;
                            acc[qi] = __builtin_fmaf(e, e, acc[qi]);
                            e = q.y - x.y;
                            acc[qi] = __builtin_fmaf(e, e, acc[qi]);
                            e = q.z - x.z;
                            acc[qi] = __builtin_fmaf(e, e, acc[qi]);
                            e = q.w - x.w;
                            acc[qi] = __builtin_fmaf(e, e, acc[qi]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int qi = 0; qi < QW; qi++) {
            if (qi >= nqg) break;
            const float d = METRIC == METRIC_IP ? -acc[qi] : acc[qi];
            const uint64_t key = valid ? (((uint64_t)order_bits(d) << 32) | (uint32_t)r) : KEY_NONE;
            sets[qi].push_lanes(key, lane);
        }
    }
    // write the slice's lists, unsorted, KEY_NONE padded
#pragma unroll
    for (int qi = 0; qi < QW; qi++) {
        if (qi >= nqg) break;
        uint64_t* out = a.partial + ((size_t)(q0 + qi) * a.parts + part) * a.k;
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int i = s * 64 + lane;
            if (i < a.k) out[i] = i < sets[qi].sz ? (uint64_t)sets[qi].key[s] : KEY_NONE;
        }
    }
}

template <int S>
__global__ void __launch_bounds__(64) gt_merge_kernel(GtArgs a) {
    const int lane = __lane_id();
    const int q = blockIdx.x;
    if (q >= a.nq) return;
    KeySet<S> set;
    set.init(a.k);
    const uint64_t* in = a.partial + (size_t)q * a.parts * a.k;
    const int total = a.parts * a.k;
    for (int b = 0; b < total; b += 64) {
        const int i = b + lane;
        const uint64_t k = i < total ? in[i] : KEY_NONE;
        set.push_lanes(k, lane);
    }
    // sorted output: repeated minimum
    for (int i = 0; i < a.k; i++) {
        uint64_t m = KEY_NONE;
#pragma unroll
        for (int s = 0; s < S; s++)
            if (s * 64 + lane < set.sz && set.key[s] < m) m = set.key[s];
        m = wave_min_u64(m);
        if (lane == 0) {
            const bool ok = m != KEY_NONE;
            a.ids_out[(size_t)q * a.k + i] = ok ? (int32_t)(uint32_t)m : -1;
            if (a.dists_out) {
                float d = INFINITY;
                if (ok) d = unorder_bits((uint32_t)(m >> 32));
                a.dists_out[(size_t)q * a.k + i] = d;
            }
        }
#pragma unroll
        for (int s = 0; s < S; s++)
            if (s * 64 + lane < set.sz && set.key[s] == m) set.key[s] = KEY_NONE;
    }
}

template <int METRIC, int S>
hipError_t launch_partial_q(const GtArgs& a, int qw, hipStream_t st) {
    dim3 grid((a.nq + qw - 1) / qw, a.parts);
    const int d4 = a.dstride / 4;
    const size_t lds = METRIC == METRIC_HAMMING ? (size_t)qw * 32 : (size_t)qw * d4 * 16;
    switch (qw) {
    case 1: hipLaunchKernelGGL((gt_partial_kernel<METRIC, S, 1>), grid, dim3(64), lds, st, a); break;
    case 2: hipLaunchKernelGGL((gt_partial_kernel<METRIC, S, 2>), grid, dim3(64), lds, st, a); break;
    case 4: hipLaunchKernelGGL((gt_partial_kernel<METRIC, S, 4>), grid, dim3(64), lds, st, a); break;
    case 8: hipLaunchKernelGGL((gt_partial_kernel<METRIC, S, 8>), grid, dim3(64), lds, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int METRIC>
hipError_t launch_partial_s(const GtArgs& a, int qw, hipStream_t st) {
    return a.k <= 64 ? launch_partial_q<METRIC, 1>(a, qw, st) : launch_partial_q<METRIC, 2>(a, qw, st);
}

}  // namespace

hipError_t launch_bruteforce(const GtArgs& a, int metric, int qw, hipStream_t st) {
    hipError_t e;
    switch (metric) {
    case METRIC_L2: e = launch_partial_s<METRIC_L2>(a, qw, st); break;
    case METRIC_IP: e = launch_partial_s<METRIC_IP>(a, qw, st); break;
    case METRIC_HAMMING: e = launch_partial_s<METRIC_HAMMING>(a, qw, st); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    if (a.k <= 64) hipLaunchKernelGGL((gt_merge_kernel<1>), dim3(a.nq), dim3(64), 0, st, a);
    else hipLaunchKernelGGL((gt_merge_kernel<2>), dim3(a.nq), dim3(64), 0, st, a);
    return hipGetLastError();
}

}  // namespace song
