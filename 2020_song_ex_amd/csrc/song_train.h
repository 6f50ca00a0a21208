// song_train.h -- internal interface of the MLP training kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace song {

// Supported shapes: (in, out) = (8, 1) stop network, (16, 1) cost network,
// (16, 64) entry network.  mse (out == 1 only): squared error on real-valued
// targets instead of the logistic loss on 0/1 labels.
int mlp_param_count(int in, int out);
hipError_t launch_mlp_grad(int in, int out, int mse, const float* X, const float* T, const int32_t* order, int64_t b0,
                           int nb, int per_wave, const float* W, float* partial, int nwaves, hipStream_t st);
hipError_t launch_mlp_adam(float* W, float* M, float* V, const float* partial, int nparts, int total, int nb, float lr,
                           float bc1, float bc2, float* loss_acc, hipStream_t st);
hipError_t launch_mlp_forward_train_layout(int in, int out, const float* X, int64_t n, const float* W, float* Y,
                                           hipStream_t st);

}  // namespace song
