// song_search_l2.hip -- instantiates the search kernels for METRIC_L2.
#include "song_search_impl.h"

namespace song {

hipError_t launch_search_l2(const SearchArgs& a, int S, int G, int qch, const LaunchShape& shape, hipStream_t st) {
    return launch_s<METRIC_L2>(a, S, G, qch, shape, st);
}

// The search kernel's inference routines on given inputs (for checking them
// against a plain FP32 MLP): stop network (8 -> 1) or entry network (16 -> 64).
__global__ void __launch_bounds__(64) stop_forward_kernel(LearnedArgs L, const float* __restrict__ X, int64_t n,
                                                          float* __restrict__ Y) {
    __shared__ float lds[LM_STOP_FLOATS];
    stage_stop_weights(L.s_weights, lds);
    __syncthreads();
    const int lane = lane_id();
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        float f[LM_F];
#pragma unroll
        for (int j = 0; j < LM_F; j++) f[j] = X[i * LM_F + j];
        const float y = stop_logit(lds, f, L.s_b3, lane);
        if (lane == 0) Y[i] = y;
    }
}

__global__ void __launch_bounds__(64) entry_forward_kernel(LearnedArgs L, const float* __restrict__ X, int64_t n,
                                                           float* __restrict__ Y) {
    const int lane = lane_id();
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        float z[LM_P];
#pragma unroll
        for (int j = 0; j < LM_P; j++) z[j] = X[i * LM_P + j];
        Y[i * LM_C + lane] = entry_scores(L, z, lane);
    }
}

hipError_t launch_learned_forward(const LearnedArgs& L, int which, const float* X, int64_t n, float* Y,
                                  hipStream_t st) {
    const int blocks = (int)(n < 4096 ? (n > 0 ? n : 1) : 4096);
    if (which == 0) hipLaunchKernelGGL(stop_forward_kernel, dim3(blocks), dim3(64), 0, st, L, X, n, Y);
    else hipLaunchKernelGGL(entry_forward_kernel, dim3(blocks), dim3(64), 0, st, L, X, n, Y);
    return hipGetLastError();
}

hipError_t launch_search(const SearchArgs& a, int metric, int S, int G, int qch, const LaunchShape& shape,
                         hipStream_t st) {
    switch (metric) {
    case METRIC_L2: return launch_search_l2(a, S, G, qch, shape, st);
    case METRIC_IP: return launch_search_ip(a, S, G, qch, shape, st);
    case METRIC_HAMMING: return launch_search_hamming(a, S, G, qch, shape, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace song
