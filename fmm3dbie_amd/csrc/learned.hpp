// Learned order selection (8 -> 64 -> 64 -> 6 MLP, FP32) on the device.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

constexpr int MLP_IN = 8, MLP_HID = 64, MLP_OUT = 6;
constexpr int MLP_NPARAM = MLP_HID * MLP_IN + MLP_HID + MLP_HID * MLP_HID + MLP_HID + MLP_OUT * MLP_HID + MLP_OUT;

// Mean softmax cross-entropy gradient over n samples (X: n x 8 row-major, y: class ids).
int mlp_grad_device(const float* params, const float* X, const int* y, int64_t n, float* grad, float* loss,
                    hipStream_t s);
// Minibatch Adam training (batch a multiple of 64, shuffled each epoch from `seed`);
// loss_hist (host, `epochs` entries) receives the last minibatch loss of every epoch.
int mlp_train_device(float* params, const float* X, const int* y, int64_t n, int epochs, int batch, float lr,
                     unsigned seed, float* loss_hist, hipStream_t s);
// Argmax class (and optionally the 6 logits) for n samples.
int mlp_predict_device(const float* params, const float* X, int64_t n, int* cls, float* logits, hipStream_t s);
