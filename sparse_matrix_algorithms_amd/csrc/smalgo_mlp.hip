// Two-layer perceptron (nin -> nh ReLU -> nout) trained and evaluated with
// hand-written HIP kernels in FP32: forward, MSE backward, Adam.  Used to map
// KPM density-of-states moments plus the target interval to POLFED filter
// parameters.  No reference counterpart.
//
// Row-major layout: X [B x nin], W1 [nin x nh], b1 [nh], H [B x nh],
// W2 [nh x nout], b2 [nout], Y [B x nout].
#include "smalgo_mlp.h"

#include <cmath>
#include <cstring>

namespace smalgo {

// Y[i, j] = act( sum_k A[i,k] W[k,j] + b[j] ), act = ReLU or identity
__global__ void dense_fwd_kernel(int B, int nin, int nout, const float* __restrict__ A,
                                 const float* __restrict__ W, const float* __restrict__ bias,
                                 float* __restrict__ Y, int relu) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * nout) return;
  const int i = idx / nout, j = idx % nout;
  float s = bias[j];
  for (int k = 0; k < nin; ++k) s += A[(size_t)i * nin + k] * W[(size_t)k * nout + j];
  Y[idx] = relu ? fmaxf(s, 0.f) : s;
}

// dY = 2 (Y - T) / (B * nout)   (gradient of the mean squared error)
__global__ void mse_grad_kernel(int n, const float* __restrict__ Y, const float* __restrict__ T,
                                float* __restrict__ dY, float scale, float* __restrict__ loss_part) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const float d = Y[idx] - T[idx];
  dY[idx] = 2.f * d * scale;
  loss_part[idx] = d * d * scale;
}

// dW[k, j] = sum_i A[i, k] dY[i, j];  db[j] = sum_i dY[i, j]
__global__ void dense_wgrad_kernel(int B, int nin, int nout, const float* __restrict__ A,
                                   const float* __restrict__ dY, float* __restrict__ dW,
                                   float* __restrict__ db) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < nin * nout) {
    const int k = idx / nout, j = idx % nout;
    float s = 0.f;
    for (int i = 0; i < B; ++i) s += A[(size_t)i * nin + k] * dY[(size_t)i * nout + j];
    dW[idx] = s;
  } else if (idx < nin * nout + nout) {
    const int j = idx - nin * nout;
    float s = 0.f;
    for (int i = 0; i < B; ++i) s += dY[(size_t)i * nout + j];
    db[j] = s;
  }
}

// dA[i, k] = (sum_j dY[i, j] W[k, j]) * (H[i,k] > 0)
__global__ void dense_agrad_relu_kernel(int B, int nin, int nout, const float* __restrict__ dY,
                                        const float* __restrict__ W, const float* __restrict__ H,
                                        float* __restrict__ dA) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * nin) return;
  const int i = idx / nin, k = idx % nin;
  float s = 0.f;
  for (int j = 0; j < nout; ++j) s += dY[(size_t)i * nout + j] * W[(size_t)k * nout + j];
  dA[idx] = H[idx] > 0.f ? s : 0.f;
}

// Adam (Kingma & Ba 2015) with bias correction
__global__ void adam_kernel(int n, float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const float gi = g[idx];
  const float mi = b1 * m[idx] + (1.f - b1) * gi;
  const float vi = b2 * v[idx] + (1.f - b2) * gi * gi;
  m[idx] = mi;
  v[idx] = vi;
  p[idx] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

static inline int nblk(long long n) { return (int)((n + 255) / 256); }

Mlp::Mlp(Ctx& c, int nin_, int nh_, int nout_, uint64_t seed) : cx(c), nin(nin_), nh(nh_), nout(nout_) {
  np = (size_t)nin * nh + nh + (size_t)nh * nout + nout;
  P.alloc(np);
  G.alloc(np);
  M.alloc(np);
  Vv.alloc(np);
  // He-uniform initialisation from splitmix64 (host), biases zero
  std::vector<float> h(np, 0.f);
  uint64_t z = seed;
  auto next = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  };
  const double a1 = std::sqrt(6.0 / nin), a2 = std::sqrt(6.0 / nh);
  for (size_t i = 0; i < (size_t)nin * nh; ++i) h[i] = (float)((2 * next() - 1) * a1);
  float* w2 = h.data() + (size_t)nin * nh + nh;
  for (size_t i = 0; i < (size_t)nh * nout; ++i) w2[i] = (float)((2 * next() - 1) * a2);
  set_params(h.data());
  SMALGO_HIP(hipMemsetAsync(M.p, 0, sizeof(float) * np, cx.stream));
  SMALGO_HIP(hipMemsetAsync(Vv.p, 0, sizeof(float) * np, cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
}

void Mlp::set_params(const float* host) {
  SMALGO_HIP(hipMemcpyAsync(P.p, host, sizeof(float) * np, hipMemcpyHostToDevice, cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
}
void Mlp::get_params(float* host) const {
  SMALGO_HIP(hipMemcpyAsync(host, P.p, sizeof(float) * np, hipMemcpyDeviceToHost, cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
}

void Mlp::ensure_batch(int B) {
  if (B <= cap) return;
  cap = B;
  X.alloc((size_t)B * nin);
  T.alloc((size_t)B * nout);
  H.alloc((size_t)B * nh);
  Y.alloc((size_t)B * nout);
  dY.alloc((size_t)B * nout);
  dH.alloc((size_t)B * nh);
  L.alloc((size_t)B * nout);
}

void Mlp::forward_dev(int B) {
  const float* W1 = P.p;
  const float* b1 = W1 + (size_t)nin * nh;
  const float* W2 = b1 + nh;
  const float* b2 = W2 + (size_t)nh * nout;
  hipLaunchKernelGGL(dense_fwd_kernel, dim3(nblk((long long)B * nh)), dim3(256), 0, cx.stream, B,
                     nin, nh, X.p, W1, b1, H.p, 1);
  hipLaunchKernelGGL(dense_fwd_kernel, dim3(nblk((long long)B * nout)), dim3(256), 0, cx.stream, B,
                     nh, nout, H.p, W2, b2, Y.p, 0);
  SMALGO_HIP(hipGetLastError());
}

float Mlp::backward_dev(int B) {
  const float* W2 = P.p + (size_t)nin * nh + nh;
  float* gW1 = G.p;
  float* gb1 = gW1 + (size_t)nin * nh;
  float* gW2 = gb1 + nh;
  float* gb2 = gW2 + (size_t)nh * nout;
  const int n = B * nout;
  hipLaunchKernelGGL(mse_grad_kernel, dim3(nblk(n)), dim3(256), 0, cx.stream, n, Y.p, T.p, dY.p,
                     1.f / (float)n, L.p);
  hipLaunchKernelGGL(dense_wgrad_kernel, dim3(nblk((long long)nh * nout + nout)), dim3(256), 0,
                     cx.stream, B, nh, nout, H.p, dY.p, gW2, gb2);
  hipLaunchKernelGGL(dense_agrad_relu_kernel, dim3(nblk((long long)B * nh)), dim3(256), 0,
                     cx.stream, B, nh, nout, dY.p, W2, H.p, dH.p);
  hipLaunchKernelGGL(dense_wgrad_kernel, dim3(nblk((long long)nin * nh + nh)), dim3(256), 0,
                     cx.stream, B, nin, nh, X.p, dH.p, gW1, gb1);
  SMALGO_HIP(hipGetLastError());
  std::vector<float> hl(n);
  SMALGO_HIP(hipMemcpyAsync(hl.data(), L.p, sizeof(float) * n, hipMemcpyDeviceToHost, cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
  double s = 0;
  for (float v : hl) s += v;
  return (float)s;
}

void Mlp::load_batch(const float* Xh, const float* Th, int B) {
  ensure_batch(B);
  SMALGO_HIP(hipMemcpyAsync(X.p, Xh, sizeof(float) * B * nin, hipMemcpyHostToDevice, cx.stream));
  if (Th)
    SMALGO_HIP(hipMemcpyAsync(T.p, Th, sizeof(float) * B * nout, hipMemcpyHostToDevice, cx.stream));
}

void Mlp::predict(const float* Xh, int B, float* Yh) {
  load_batch(Xh, nullptr, B);
  forward_dev(B);
  SMALGO_HIP(hipMemcpyAsync(Yh, Y.p, sizeof(float) * B * nout, hipMemcpyDeviceToHost, cx.stream));
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
}

float Mlp::gradient(const float* Xh, const float* Th, int B, float* gh) {
  load_batch(Xh, Th, B);
  forward_dev(B);
  const float loss = backward_dev(B);
  if (gh) {
    SMALGO_HIP(hipMemcpyAsync(gh, G.p, sizeof(float) * np, hipMemcpyDeviceToHost, cx.stream));
    SMALGO_HIP(hipStreamSynchronize(cx.stream));
  }
  return loss;
}

float Mlp::train(const float* Xh, const float* Th, int B, int epochs, float lr, float* loss_hist) {
  load_batch(Xh, Th, B);
  float loss = 0.f;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int e = 0; e < epochs; ++e) {
    forward_dev(B);
    loss = backward_dev(B);
    if (loss_hist) loss_hist[e] = loss;
    ++step;
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(nblk((long long)np)), dim3(256), 0, cx.stream, (int)np,
                       P.p, G.p, M.p, Vv.p, lr, b1, b2, eps, c1, c2);
    SMALGO_HIP(hipGetLastError());
  }
  SMALGO_HIP(hipStreamSynchronize(cx.stream));
  return loss;
}

}  // namespace smalgo
