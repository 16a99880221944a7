// Learned order selection: an 8 -> 64 -> 64 -> 6 MLP (ReLU, softmax cross-entropy)
// trained on the GPU to predict each adaptive pair's rule class (the XG degree of least
// cost, labelled from the adaptive integrator's own runs).
// Parameters (FP32, packed): W1[64][8] b1[64] W2[64][64] b2[64] W3[6][64] b3[6].
//
// Training step = k_mlp_grad (64 samples per workgroup: forward, backward, per-workgroup
// gradient partials, no atomics) + k_mlp_reduce_adam (sum partials in a fixed order,
// Adam update).  Inference = k_mlp_predict (argmax class).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "learned.hpp"

namespace {

constexpr int NI = MLP_IN, NH = MLP_HID, NO = MLP_OUT;
constexpr int OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NO * NH, NPAR = OB3 + NO;
static_assert(NPAR == MLP_NPARAM, "parameter count");
constexpr int SB = 64;  // samples per workgroup

__device__ __forceinline__ void forward1(const float* P, const float* x, float* h1, float* h2, float* z) {
  for (int i = 0; i < NH; i++) {
    float a = P[OB1 + i];
#pragma unroll
    for (int k = 0; k < NI; k++) a = fmaf(P[OW1 + i * NI + k], x[k], a);
    h1[i] = a > 0.f ? a : 0.f;
  }
  for (int i = 0; i < NH; i++) {
    float a = P[OB2 + i];
    for (int k = 0; k < NH; k++) a = fmaf(P[OW2 + i * NH + k], h1[k], a);
    h2[i] = a > 0.f ? a : 0.f;
  }
  for (int i = 0; i < NO; i++) {
    float a = P[OB3 + i];
    for (int k = 0; k < NH; k++) a = fmaf(P[OW3 + i * NH + k], h2[k], a);
    z[i] = a;
  }
}

// grads of the mean cross-entropy over `batch` samples starting at `start`; one
// workgroup of 64 threads handles 64 samples and writes its partial gradient.
__global__ void __launch_bounds__(SB) k_mlp_grad(const float* __restrict__ params, const float* __restrict__ X,
                                                 const int* __restrict__ y, const int* __restrict__ perm, int64_t start,
                                                 int batch, float inv_batch, float* __restrict__ partial,
                                                 float* __restrict__ loss_part) {
  __shared__ float P[NPAR];
  __shared__ float sx[SB][NI];
  __shared__ float sh1[SB][NH], sh2[SB][NH], sd2[SB][NH], sd1[SB][NH];
  __shared__ float sdz[SB][NO];
  __shared__ float sl[SB];
  const int t = threadIdx.x;
  for (int i = t; i < NPAR; i += SB) P[i] = params[i];
  __syncthreads();
  const int s = blockIdx.x * SB + t;
  const bool act = s < batch;
  float x[NI], h1[NH], h2[NH], z[NO];
  int lab = 0;
  if (act) {
    const int64_t idx = perm ? perm[start + s] : start + s;
#pragma unroll
    for (int k = 0; k < NI; k++) x[k] = X[idx * NI + k];
    lab = y[idx];
  } else {
#pragma unroll
    for (int k = 0; k < NI; k++) x[k] = 0.f;
  }
  forward1(P, x, h1, h2, z);
  float mx = z[0];
  for (int i = 1; i < NO; i++) mx = fmaxf(mx, z[i]);
  float se = 0.f;
  for (int i = 0; i < NO; i++) se += expf(z[i] - mx);
  float dz[NO];
  for (int i = 0; i < NO; i++) {
    const float pr = expf(z[i] - mx) / se;
    dz[i] = act ? (pr - (i == lab ? 1.f : 0.f)) * inv_batch : 0.f;
  }
  sl[t] = act ? (logf(se) + mx - z[lab]) * inv_batch : 0.f;
  // dh2 = W3^T dz .* relu'(h2); dh1 = W2^T dh2 .* relu'(h1)
  for (int k = 0; k < NH; k++) {
    float a = 0.f;
    for (int i = 0; i < NO; i++) a = fmaf(P[OW3 + i * NH + k], dz[i], a);
    sd2[t][k] = h2[k] > 0.f ? a : 0.f;
  }
  for (int k = 0; k < NH; k++) {
    float a = 0.f;
    for (int i = 0; i < NH; i++) a = fmaf(P[OW2 + i * NH + k], sd2[t][i], a);
    sd1[t][k] = h1[k] > 0.f ? a : 0.f;
  }
  for (int k = 0; k < NI; k++) sx[t][k] = x[k];
  for (int k = 0; k < NH; k++) {
    sh1[t][k] = h1[k];
    sh2[t][k] = h2[k];
  }
  for (int i = 0; i < NO; i++) sdz[t][i] = dz[i];
  __syncthreads();
  // thread t owns row t of W1/W2 (and of W3 for t < 6) and the matching biases
  float* g = partial + (size_t)blockIdx.x * NPAR;
  {
    float gb1 = 0.f, gb2 = 0.f;
    for (int ss = 0; ss < SB; ss++) {
      gb1 += sd1[ss][t];
      gb2 += sd2[ss][t];
    }
    g[OB1 + t] = gb1;
    g[OB2 + t] = gb2;
    for (int k = 0; k < NI; k++) {
      float a = 0.f;
      for (int ss = 0; ss < SB; ss++) a = fmaf(sd1[ss][t], sx[ss][k], a);
      g[OW1 + t * NI + k] = a;
    }
    for (int k = 0; k < NH; k++) {
      float a = 0.f;
      for (int ss = 0; ss < SB; ss++) a = fmaf(sd2[ss][t], sh1[ss][k], a);
      g[OW2 + t * NH + k] = a;
    }
    if (t < NO) {
      float gb3 = 0.f;
      for (int ss = 0; ss < SB; ss++) gb3 += sdz[ss][t];
      g[OB3 + t] = gb3;
      for (int k = 0; k < NH; k++) {
        float a = 0.f;
        for (int ss = 0; ss < SB; ss++) a = fmaf(sdz[ss][t], sh2[ss][k], a);
        g[OW3 + t * NH + k] = a;
      }
    }
  }
  __syncthreads();
  if (t == 0) {
    float a = 0.f;
    for (int ss = 0; ss < SB; ss++) a += sl[ss];
    loss_part[blockIdx.x] = a;
  }
}

// grad = sum of the nb partials (fixed order); optional Adam step
__global__ void k_mlp_reduce_adam(int nb, const float* __restrict__ partial, float* __restrict__ grad,
                                  float* __restrict__ params, float* __restrict__ m, float* __restrict__ v, float lr,
                                  float b1, float b2, float eps, float bc1, float bc2, int do_update) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPAR) return;
  float gs = 0.f;
  for (int b = 0; b < nb; b++) gs += partial[(size_t)b * NPAR + i];
  grad[i] = gs;
  if (do_update) {
    const float mi = b1 * m[i] + (1.f - b1) * gs;
    const float vi = b2 * v[i] + (1.f - b2) * gs * gs;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  }
}

__global__ void k_mlp_predict(int64_t n, const float* __restrict__ params, const float* __restrict__ X,
                              int* __restrict__ cls, float* __restrict__ logits) {
  __shared__ float P[NPAR];
  for (int i = threadIdx.x; i < NPAR; i += blockDim.x) P[i] = params[i];
  __syncthreads();
  const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[NI], h1[NH], h2[NH], z[NO];
#pragma unroll
  for (int k = 0; k < NI; k++) x[k] = X[s * NI + k];
  forward1(P, x, h1, h2, z);
  int best = 0;
  for (int i = 1; i < NO; i++)
    if (z[i] > z[best]) best = i;
  if (cls) cls[s] = best;
  if (logits)
    for (int i = 0; i < NO; i++) logits[s * NO + i] = z[i];
}

}  // namespace

int mlp_grad_device(const float* params, const float* X, const int* y, int64_t n, float* grad, float* loss,
                    hipStream_t s) {
  const int nb = (int)((n + SB - 1) / SB);
  float *part = nullptr, *lp = nullptr;
  if (hipMalloc((void**)&part, sizeof(float) * (size_t)nb * NPAR) != hipSuccess) return -2;
  if (hipMalloc((void**)&lp, sizeof(float) * nb) != hipSuccess) return -2;
  k_mlp_grad<<<nb, SB, 0, s>>>(params, X, y, nullptr, 0, (int)n, 1.0f / (float)n, part, lp);
  k_mlp_reduce_adam<<<(NPAR + 255) / 256, 256, 0, s>>>(nb, part, grad, nullptr, nullptr, nullptr, 0, 0, 0, 0, 1, 1,
                                                       0);
  std::vector<float> hl(nb);
  (void)hipMemcpyAsync(hl.data(), lp, sizeof(float) * nb, hipMemcpyDeviceToHost, s);
  (void)hipStreamSynchronize(s);
  float L = 0.f;
  for (float v : hl) L += v;
  if (loss) *loss = L;
  (void)hipFree(part);
  (void)hipFree(lp);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mlp_train_device(float* params, const float* X, const int* y, int64_t n, int epochs, int batch, float lr,
                     unsigned seed, float* loss_hist, hipStream_t s) {
  if (batch % SB != 0 || batch <= 0) return -6;
  const int nb = batch / SB;
  float *part = nullptr, *lp = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  int* perm = nullptr;
  if (hipMalloc((void**)&part, sizeof(float) * (size_t)nb * NPAR) != hipSuccess) return -2;
  if (hipMalloc((void**)&lp, sizeof(float) * nb) != hipSuccess) return -2;
  if (hipMalloc((void**)&grad, sizeof(float) * NPAR) != hipSuccess) return -2;
  if (hipMalloc((void**)&m, sizeof(float) * NPAR) != hipSuccess) return -2;
  if (hipMalloc((void**)&v, sizeof(float) * NPAR) != hipSuccess) return -2;
  if (hipMalloc((void**)&perm, sizeof(int) * n) != hipSuccess) return -2;
  (void)hipMemsetAsync(m, 0, sizeof(float) * NPAR, s);
  (void)hipMemsetAsync(v, 0, sizeof(float) * NPAR, s);
  std::vector<int> hp(n);
  unsigned st = seed ? seed : 1u;
  const float b1 = 0.9f, b2 = 0.999f;
  int step = 0;
  const int64_t nsteps_ep = n / batch;
  for (int ep = 0; ep < epochs; ep++) {
    (void)hipStreamSynchronize(s);  // the previous epoch's upload of hp has completed
    for (int64_t i = 0; i < n; i++) hp[i] = (int)i;
    for (int64_t i = n - 1; i > 0; i--) {  // Fisher-Yates with an LCG
      st = st * 1664525u + 1013904223u;
      const int64_t j = st % (uint64_t)(i + 1);
      std::swap(hp[i], hp[j]);
    }
    (void)hipMemcpyAsync(perm, hp.data(), sizeof(int) * n, hipMemcpyHostToDevice, s);
    const float lr_ep = lr * (ep < epochs / 2 ? 1.0f : (ep < 3 * epochs / 4 ? 0.3f : 0.1f));
    double eloss = 0.0;
    std::vector<float> hl(nb);
    for (int64_t k = 0; k < nsteps_ep; k++) {
      step++;
      k_mlp_grad<<<nb, SB, 0, s>>>(params, X, y, perm, k * batch, batch, 1.0f / batch, part, lp);
      const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
      k_mlp_reduce_adam<<<(NPAR + 255) / 256, 256, 0, s>>>(nb, part, grad, params, m, v, lr_ep, b1, b2, 1e-8f, bc1,
                                                           bc2, 1);
      if (loss_hist && k == nsteps_ep - 1) {
        (void)hipMemcpyAsync(hl.data(), lp, sizeof(float) * nb, hipMemcpyDeviceToHost, s);
        (void)hipStreamSynchronize(s);
        for (float x : hl) eloss += x;
      }
    }
    if (loss_hist) loss_hist[ep] = (float)eloss;
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(part);
  (void)hipFree(lp);
  (void)hipFree(grad);
  (void)hipFree(m);
  (void)hipFree(v);
  (void)hipFree(perm);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mlp_predict_device(const float* params, const float* X, int64_t n, int* cls, float* logits, hipStream_t s) {
  if (n <= 0) return 0;
  k_mlp_predict<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, params, X, cls, logits);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
