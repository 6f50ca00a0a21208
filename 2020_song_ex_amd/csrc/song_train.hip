// song_train.hip -- training of the 2 x 64 ReLU MLPs of the learned search
// (entry and stop networks) on the device, with Adam.
//
// Weights (flat, row-major): w1 [64 x IN], b1 [64], w2 [64 x 64], b2 [64],
// w3 [OUT x 64], b3 [OUT].  One wave handles a run of samples of the
// minibatch, lane j standing for hidden unit j (and output unit j); it keeps
// its gradient share in registers and writes one partial gradient vector per
// wave.  A second kernel sums the partials in a fixed order and takes the Adam
// step, so training is deterministic: no atomics anywhere.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "song_train.h"

namespace song {

namespace {

constexpr int H = 64;
constexpr int PAD = H + 1;  // LDS row stride: lane j reading row j hits bank (j + i) % 32

__device__ __forceinline__ float rl(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ float xsum(float v) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__device__ __forceinline__ float xmax(float v) {
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
    return v;
}

struct Off {
    int w1, b1, w2, b2, w3, b3, total;
};

__host__ __device__ inline Off offsets(int in, int out) {
    Off o;
    o.w1 = 0;
    o.b1 = H * in;
    o.w2 = o.b1 + H;
    o.b2 = o.w2 + H * H;
    o.w3 = o.b2 + H;
    o.b3 = o.w3 + out * H;
    o.total = o.b3 + out;
    return o;
}

template <int IN, int OUT>
struct Lds {
    float w1[H * (IN + 1)];
    float b1[H];
    float w2[H * PAD];
    float b2[H];
    float w3[OUT * PAD];
    float b3[OUT];
};

template <int IN, int OUT>
__device__ void stage(const float* W, Lds<IN, OUT>& s) {
    const Off o = offsets(IN, OUT);
    const int lane = threadIdx.x;
    for (int i = lane; i < H * IN; i += 64) s.w1[(i / IN) * (IN + 1) + i % IN] = W[o.w1 + i];
    for (int i = lane; i < H * H; i += 64) s.w2[(i / H) * PAD + i % H] = W[o.w2 + i];
    for (int i = lane; i < OUT * H; i += 64) s.w3[(i / H) * PAD + i % H] = W[o.w3 + i];
    s.b1[lane] = W[o.b1 + lane];
    s.b2[lane] = W[o.b2 + lane];
    if (lane < OUT) s.b3[lane] = W[o.b3 + lane];
    __syncthreads();
}

// Forward of one sample (x wave-uniform); lane j: pre-activations of unit j,
// output c = lane for lane < OUT.
template <int IN, int OUT>
__device__ __forceinline__ void forward(const Lds<IN, OUT>& s, const float (&x)[IN], int lane, float& pre1, float& h1,
                                        float& pre2, float& h2, float& o) {
    pre1 = s.b1[lane];
#pragma unroll
    for (int i = 0; i < IN; i++) pre1 = __builtin_fmaf(s.w1[lane * (IN + 1) + i], x[i], pre1);
    h1 = fmaxf(pre1, 0.f);
    pre2 = s.b2[lane];
#pragma unroll
    for (int i = 0; i < H; i++) pre2 = __builtin_fmaf(s.w2[lane * PAD + i], rl(h1, i), pre2);
    h2 = fmaxf(pre2, 0.f);
    o = 0.f;
    if (lane < OUT) {
        o = s.b3[lane];
#pragma unroll
        for (int j = 0; j < H; j++) o = __builtin_fmaf(s.w3[lane * PAD + j], rl(h2, j), o);
    }
}

template <int IN, int OUT>
__global__ void __launch_bounds__(64) mlp_grad_kernel(const float* __restrict__ X, const float* __restrict__ T,
                                                      const int32_t* __restrict__ order, int64_t b0, int nb,
                                                      int per_wave, int mse, const float* __restrict__ W,
                                                      float* __restrict__ partial) {
    __shared__ Lds<IN, OUT> s;
    stage<IN, OUT>(W, s);
    const int lane = threadIdx.x;
    const Off off = offsets(IN, OUT);
    float gw1[IN], gw2[H], gw3[OUT];
#pragma unroll
    for (int i = 0; i < IN; i++) gw1[i] = 0.f;
#pragma unroll
    for (int i = 0; i < H; i++) gw2[i] = 0.f;
#pragma unroll
    for (int c = 0; c < OUT; c++) gw3[c] = 0.f;
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, loss = 0.f;
    const int first = blockIdx.x * per_wave;
    for (int t = 0; t < per_wave && first + t < nb; t++) {
        const int64_t idx = order[b0 + first + t];
        float x[IN];
#pragma unroll
        for (int i = 0; i < IN; i++) x[i] = X[idx * IN + i];
        float pre1, h1, pre2, h2, o;
        forward<IN, OUT>(s, x, lane, pre1, h1, pre2, h2, o);
        float g = 0.f;  // dLoss/d output (lane c)
        if constexpr (OUT == 1) {
            const float z = rl(o, 0), y = T[idx];
            if (mse) {  // squared error / 2 on a real-valued target
                if (lane == 0) g = z - y;
                loss += 0.5f * (z - y) * (z - y);
            } else {    // logistic loss on a 0/1 label
                const float p = 1.f / (1.f + __expf(-z));
                if (lane == 0) g = p - y;
                loss += fmaxf(z, 0.f) - z * y + __logf(1.f + __expf(-fabsf(z)));
            }
        } else {
            const int label = (int)T[idx];
            const float m = xmax(lane < OUT ? o : -INFINITY);
            const float e = lane < OUT ? __expf(o - m) : 0.f;
            const float Z = xsum(e);
            const float p = e / Z;
            if (lane < OUT) g = p - (lane == label ? 1.f : 0.f);
            loss += __logf(Z) - (rl(o, label) - m);
        }
        gb3 += g;
        float dh2 = 0.f;
#pragma unroll
        for (int c = 0; c < OUT; c++) {
            const float gc = rl(g, c);
            gw3[c] = __builtin_fmaf(gc, h2, gw3[c]);
            dh2 = __builtin_fmaf(s.w3[c * PAD + lane], gc, dh2);
        }
        const float dpre2 = pre2 > 0.f ? dh2 : 0.f;
        gb2 += dpre2;
#pragma unroll
        for (int i = 0; i < H; i++) gw2[i] = __builtin_fmaf(dpre2, rl(h1, i), gw2[i]);
        float dh1 = 0.f;
#pragma unroll
        for (int j = 0; j < H; j++) dh1 = __builtin_fmaf(s.w2[j * PAD + lane], rl(dpre2, j), dh1);
        const float dpre1 = pre1 > 0.f ? dh1 : 0.f;
        gb1 += dpre1;
#pragma unroll
        for (int i = 0; i < IN; i++) gw1[i] = __builtin_fmaf(dpre1, x[i], gw1[i]);
    }
    float* out = partial + (size_t)blockIdx.x * (off.total + 1);
#pragma unroll
    for (int i = 0; i < IN; i++) out[off.w1 + lane * IN + i] = gw1[i];
    out[off.b1 + lane] = gb1;
#pragma unroll
    for (int i = 0; i < H; i++) out[off.w2 + lane * H + i] = gw2[i];
    out[off.b2 + lane] = gb2;
#pragma unroll
    for (int c = 0; c < OUT; c++) out[off.w3 + c * H + lane] = gw3[c];
    if (lane < OUT) out[off.b3 + lane] = gb3;
    if (lane == 0) out[off.total] = loss;
}

// Sum the wave partials (fixed order), Adam step.  The thread of parameter 0
// also accumulates the minibatch loss.
__global__ void mlp_adam_kernel(float* __restrict__ W, float* __restrict__ M, float* __restrict__ V,
                                const float* __restrict__ partial, int nparts, int total, int nb, float lr,
                                float bc1, float bc2, float* __restrict__ loss_acc) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < total) {
        float g = 0.f;
        for (int w = 0; w < nparts; w++) g += partial[(size_t)w * (total + 1) + p];
        g /= (float)nb;
        const float m = 0.9f * M[p] + 0.1f * g;
        const float v = 0.999f * V[p] + 0.001f * g * g;
        M[p] = m;
        V[p] = v;
        W[p] -= lr * (m / bc1) / (sqrtf(v / bc2) + 1e-8f);
    }
    if (p == 0) {
        float l = 0.f;
        for (int w = 0; w < nparts; w++) l += partial[(size_t)w * (total + 1) + total];
        *loss_acc += l;
    }
}

template <int IN, int OUT>
__global__ void __launch_bounds__(64) mlp_forward_train_layout(const float* __restrict__ X, int64_t n,
                                                                const float* __restrict__ W, float* __restrict__ Y) {
    __shared__ Lds<IN, OUT> s;
    stage<IN, OUT>(W, s);
    const int lane = threadIdx.x;
    for (int64_t idx = blockIdx.x; idx < n; idx += gridDim.x) {
        float x[IN];
#pragma unroll
        for (int i = 0; i < IN; i++) x[i] = X[idx * IN + i];
        float pre1, h1, pre2, h2, o;
        forward<IN, OUT>(s, x, lane, pre1, h1, pre2, h2, o);
        if (lane < OUT) Y[idx * OUT + lane] = o;
    }
}

}  // namespace

int mlp_param_count(int in, int out) { return offsets(in, out).total; }

hipError_t launch_mlp_grad(int in, int out, int mse, const float* X, const float* T, const int32_t* order, int64_t b0,
                           int nb, int per_wave, const float* W, float* partial, int nwaves, hipStream_t st) {
    if (in == 8 && out == 1)
        hipLaunchKernelGGL((mlp_grad_kernel<8, 1>), dim3(nwaves), dim3(64), 0, st, X, T, order, b0, nb, per_wave, mse,
                           W, partial);
    else if (in == 16 && out == 1)
        hipLaunchKernelGGL((mlp_grad_kernel<16, 1>), dim3(nwaves), dim3(64), 0, st, X, T, order, b0, nb, per_wave, mse,
                           W, partial);
    else if (in == 16 && out == 64)
        hipLaunchKernelGGL((mlp_grad_kernel<16, 64>), dim3(nwaves), dim3(64), 0, st, X, T, order, b0, nb, per_wave, 0,
                           W, partial);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float* W, float* M, float* V, const float* partial, int nparts, int total, int nb, float lr,
                           float bc1, float bc2, float* loss_acc, hipStream_t st) {
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((total + 255) / 256), dim3(256), 0, st, W, M, V, partial, nparts, total,
                       nb, lr, bc1, bc2, loss_acc);
    return hipGetLastError();
}

hipError_t launch_mlp_forward_train_layout(int in, int out, const float* X, int64_t n, const float* W, float* Y,
                                           hipStream_t st) {
    const int blocks = (int)(n < 4096 ? (n > 0 ? n : 1) : 4096);
    if (in == 8 && out == 1)
        hipLaunchKernelGGL((mlp_forward_train_layout<8, 1>), dim3(blocks), dim3(64), 0, st, X, n, W, Y);
    else if (in == 16 && out == 1)
        hipLaunchKernelGGL((mlp_forward_train_layout<16, 1>), dim3(blocks), dim3(64), 0, st, X, n, W, Y);
    else if (in == 16 && out == 64)
        hipLaunchKernelGGL((mlp_forward_train_layout<16, 64>), dim3(blocks), dim3(64), 0, st, X, n, W, Y);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace song
