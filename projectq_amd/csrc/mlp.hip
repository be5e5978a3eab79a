// HIP kernels of the planner's MLP cost model (see mlp.hpp): row-parallel forward and backward
// passes, per-weight gradient reductions in a fixed row order (deterministic), and Adam.
#include "mlp.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>

#include "common.hpp"

namespace qs {
namespace mlp {
namespace {

constexpr int kThreads = 256;

// Forward pass for rows idx[r] (idx == nullptr: row r).  Stores post-activation hidden layers.
__global__ __launch_bounds__(kThreads) void k_forward(const float *__restrict__ w, const float *__restrict__ X,
                                                      const int *__restrict__ idx, int rows, float *H1, float *H2,
                                                      float *Y) {
  __shared__ float W[kWeights];
  for (int i = threadIdx.x; i < kWeights; i += kThreads) W[i] = w[i];
  __syncthreads();
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= rows) return;
  const int src = idx ? idx[r] : r;
  float x[kIn];
#pragma unroll
  for (int k = 0; k < kIn; ++k) x[k] = X[(size_t)src * kIn + k];
  float h1[kHidden];
#pragma unroll
  for (int j = 0; j < kHidden; ++j) {
    float a = W[kB1 + j];
#pragma unroll
    for (int k = 0; k < kIn; ++k) a = fmaf(W[kW1 + j * kIn + k], x[k], a);
    h1[j] = fmaxf(a, 0.f);
    if (H1) H1[(size_t)r * kHidden + j] = h1[j];
  }
  float y = W[kB3];
  for (int j = 0; j < kHidden; ++j) {
    float a = W[kB2 + j];
#pragma unroll
    for (int k = 0; k < kHidden; ++k) a = fmaf(W[kW2 + j * kHidden + k], h1[k], a);
    const float h2 = fmaxf(a, 0.f);
    if (H2) H2[(size_t)r * kHidden + j] = h2;
    y = fmaf(W[kW3 + j], h2, y);
  }
  Y[r] = y;
}

// Per-row backward: dY = 2 (y - t) / rows, dH2, dH1 (gradients w.r.t. the pre-activations).
__global__ __launch_bounds__(kThreads) void k_backward_rows(const float *__restrict__ w, const float *__restrict__ T,
                                                            const int *__restrict__ idx, int rows,
                                                            const float *__restrict__ H1, const float *__restrict__ H2,
                                                            const float *__restrict__ Y, float *dY, float *dH1,
                                                            float *dH2, float *err) {
  __shared__ float W[kWeights];
  for (int i = threadIdx.x; i < kWeights; i += kThreads) W[i] = w[i];
  __syncthreads();
  const int r = blockIdx.x * kThreads + threadIdx.x;
  if (r >= rows) return;
  const int src = idx ? idx[r] : r;
  const float e = Y[r] - T[src];
  err[r] = e * e;
  const float dy = 2.f * e / (float)rows;
  dY[r] = dy;
  float d2[kHidden];
#pragma unroll
  for (int j = 0; j < kHidden; ++j) {
    d2[j] = H2[(size_t)r * kHidden + j] > 0.f ? dy * W[kW3 + j] : 0.f;
    dH2[(size_t)r * kHidden + j] = d2[j];
  }
  for (int k = 0; k < kHidden; ++k) {
    float a = 0.f;
#pragma unroll
    for (int j = 0; j < kHidden; ++j) a = fmaf(W[kW2 + j * kHidden + k], d2[j], a);
    dH1[(size_t)r * kHidden + k] = H1[(size_t)r * kHidden + k] > 0.f ? a : 0.f;
  }
}

// One thread per weight: sum over rows in order.
__global__ __launch_bounds__(kThreads) void k_weight_grads(const float *__restrict__ X, const int *__restrict__ idx,
                                                           int rows, const float *__restrict__ H1,
                                                           const float *__restrict__ H2,
                                                           const float *__restrict__ dY,
                                                           const float *__restrict__ dH1,
                                                           const float *__restrict__ dH2, float *G) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= kWeights) return;
  float g = 0.f;
  if (i < kB1) {
    const int j = i / kIn, k = i % kIn;
    for (int r = 0; r < rows; ++r) g = fmaf(dH1[(size_t)r * kHidden + j], X[(size_t)(idx ? idx[r] : r) * kIn + k], g);
  } else if (i < kW2) {
    const int j = i - kB1;
    for (int r = 0; r < rows; ++r) g += dH1[(size_t)r * kHidden + j];
  } else if (i < kB2) {
    const int j = (i - kW2) / kHidden, k = (i - kW2) % kHidden;
    for (int r = 0; r < rows; ++r) g = fmaf(dH2[(size_t)r * kHidden + j], H1[(size_t)r * kHidden + k], g);
  } else if (i < kW3) {
    const int j = i - kB2;
    for (int r = 0; r < rows; ++r) g += dH2[(size_t)r * kHidden + j];
  } else if (i < kB3) {
    const int k = i - kW3;
    for (int r = 0; r < rows; ++r) g = fmaf(dY[r], H2[(size_t)r * kHidden + k], g);
  } else {
    for (int r = 0; r < rows; ++r) g += dY[r];
  }
  G[i] = g;
}

__global__ void k_mean(const float *__restrict__ v, int n, float *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += v[i];
  *out = (float)(s / n);
}

__global__ __launch_bounds__(kThreads) void k_adam(float *w, const float *__restrict__ g, float *m, float *v, float lr,
                                                   float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= kWeights) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline int blocks_for(size_t n) { return (int)((n + kThreads - 1) / kThreads); }

struct Scratch {
  float *H1 = nullptr, *H2 = nullptr, *Y = nullptr, *dY = nullptr, *dH1 = nullptr, *dH2 = nullptr, *err = nullptr;
  explicit Scratch(size_t rows) {
    QS_HIP(hipMalloc(&H1, rows * kHidden * sizeof(float)));
    QS_HIP(hipMalloc(&H2, rows * kHidden * sizeof(float)));
    QS_HIP(hipMalloc(&dH1, rows * kHidden * sizeof(float)));
    QS_HIP(hipMalloc(&dH2, rows * kHidden * sizeof(float)));
    QS_HIP(hipMalloc(&Y, rows * sizeof(float)));
    QS_HIP(hipMalloc(&dY, rows * sizeof(float)));
    QS_HIP(hipMalloc(&err, rows * sizeof(float)));
  }
  ~Scratch() {
    float *all[] = {H1, H2, Y, dY, dH1, dH2, err};
    for (float *p : all) (void)hipFree(p);
  }
};

template <class T>
T *dev_copy(const T *h, size_t n) {
  T *d = nullptr;
  QS_HIP(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(T)));
  if (n) QS_HIP(hipMemcpy(d, h, n * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

// loss and gradient of rows idx[0..rows) (device pointers)
void device_loss_grad(const float *dw, const float *dX, const float *dT, const int *didx, int rows, Scratch &s,
                      float *dG, float *dloss, hipStream_t st) {
  hipLaunchKernelGGL(k_forward, dim3(blocks_for(rows)), dim3(kThreads), 0, st, dw, dX, didx, rows, s.H1, s.H2, s.Y);
  QS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_backward_rows, dim3(blocks_for(rows)), dim3(kThreads), 0, st, dw, dT, didx, rows, s.H1, s.H2,
                     s.Y, s.dY, s.dH1, s.dH2, s.err);
  QS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_weight_grads, dim3(blocks_for(kWeights)), dim3(kThreads), 0, st, dX, didx, rows, s.H1, s.H2,
                     s.dY, s.dH1, s.dH2, dG);
  QS_HIP(hipGetLastError());
  if (dloss) {
    hipLaunchKernelGGL(k_mean, dim3(1), dim3(64), 0, st, s.err, rows, dloss);
    QS_HIP(hipGetLastError());
  }
}

}  // namespace

void predict(const float *w, const float *X, size_t rows, float *y) {
  if (rows == 0) return;
  Model m(std::vector<float>(w, w + kWeights));
  std::vector<float> out = m.predict(std::vector<float>(X, X + rows * kIn), rows, nullptr);
  std::copy(out.begin(), out.end(), y);
}

void loss_grad(const float *w, const float *X, const float *T, size_t rows, float *loss, float *grad) {
  if (rows == 0) throw ValueError("mlp: no rows");
  float *dw = dev_copy(w, kWeights), *dX = dev_copy(X, rows * kIn), *dT = dev_copy(T, rows);
  float *dG = nullptr, *dl = nullptr;
  try {
    QS_HIP(hipMalloc(&dG, kWeights * sizeof(float)));
    QS_HIP(hipMalloc(&dl, sizeof(float)));
    Scratch s(rows);
    device_loss_grad(dw, dX, dT, nullptr, (int)rows, s, dG, dl, nullptr);
    QS_HIP(hipMemcpy(grad, dG, kWeights * sizeof(float), hipMemcpyDeviceToHost));
    QS_HIP(hipMemcpy(loss, dl, sizeof(float), hipMemcpyDeviceToHost));
  } catch (...) {
    (void)hipFree(dw), (void)hipFree(dX), (void)hipFree(dT), (void)hipFree(dG), (void)hipFree(dl);
    throw;
  }
  (void)hipFree(dw), (void)hipFree(dX), (void)hipFree(dT), (void)hipFree(dG), (void)hipFree(dl);
}

float train(const float *X, const float *T, size_t rows, int epochs, int batch, float lr, uint32_t seed,
            const float *w_init, float *w_out) {
  if (rows == 0 || batch <= 0 || epochs < 0) throw ValueError("mlp: bad training arguments");
  std::mt19937 gen(seed);
  std::vector<float> w(kWeights, 0.f);
  if (w_init) {
    std::copy(w_init, w_init + kWeights, w.begin());
  } else {
    std::normal_distribution<float> nd(0.f, 1.f);
    for (int i = kW1; i < kB1; ++i) w[i] = nd(gen) * std::sqrt(2.f / kIn);
    for (int i = kW2; i < kB2; ++i) w[i] = nd(gen) * std::sqrt(2.f / kHidden);
    for (int i = kW3; i < kB3; ++i) w[i] = nd(gen) * std::sqrt(1.f / kHidden);
  }
  const int B = (int)std::min<size_t>(rows, (size_t)batch);
  float *dw = dev_copy(w.data(), kWeights), *dX = dev_copy(X, rows * kIn), *dT = dev_copy(T, rows);
  std::vector<float> zeros(kWeights, 0.f);
  float *dm = dev_copy(zeros.data(), kWeights), *dv = dev_copy(zeros.data(), kWeights);
  float *dG = nullptr, *dl = nullptr;
  int *didx = nullptr;
  float final_loss = 0.f;
  try {
    QS_HIP(hipMalloc(&dG, kWeights * sizeof(float)));
    QS_HIP(hipMalloc(&dl, sizeof(float)));
    QS_HIP(hipMalloc(&didx, rows * sizeof(int)));
    Scratch s(rows);
    std::vector<int> perm(rows);
    std::iota(perm.begin(), perm.end(), 0);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    int step = 0;
    for (int e = 0; e < epochs; ++e) {
      std::shuffle(perm.begin(), perm.end(), gen);
      QS_HIP(hipMemcpy(didx, perm.data(), rows * sizeof(int), hipMemcpyHostToDevice));
      for (size_t start = 0; start + B <= rows; start += B) {
        device_loss_grad(dw, dX, dT, didx + start, B, s, dG, nullptr, nullptr);
        ++step;
        const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(kWeights)), dim3(kThreads), 0, nullptr, dw, dG, dm, dv, lr, b1, b2,
                           eps, c1, c2);
        QS_HIP(hipGetLastError());
      }
    }
    device_loss_grad(dw, dX, dT, nullptr, (int)rows, s, dG, dl, nullptr);
    QS_HIP(hipMemcpy(&final_loss, dl, sizeof(float), hipMemcpyDeviceToHost));
    QS_HIP(hipMemcpy(w_out, dw, kWeights * sizeof(float), hipMemcpyDeviceToHost));
  } catch (...) {
    void *all[] = {dw, dX, dT, dm, dv, dG, dl, didx};
    for (void *p : all) (void)hipFree(p);
    throw;
  }
  void *all[] = {dw, dX, dT, dm, dv, dG, dl, didx};
  for (void *p : all) (void)hipFree(p);
  return final_loss;
}

Model::Model(const std::vector<float> &w) {
  if (w.size() != (size_t)kWeights) throw ValueError("mlp: wrong number of weights");
  d_w_ = dev_copy(w.data(), w.size());
}

Model::~Model() {
  (void)hipFree(d_w_);
  (void)hipFree(d_x_);
  (void)hipFree(d_y_);
}

std::vector<float> Model::predict(const std::vector<float> &X, size_t rows, hipStream_t s) {
  std::vector<float> y(rows);
  if (rows == 0) return y;
  if (rows > cap_) {
    (void)hipFree(d_x_);
    (void)hipFree(d_y_);
    d_x_ = d_y_ = nullptr;
    cap_ = 0;
    QS_HIP(hipMalloc(&d_x_, rows * kIn * sizeof(float)));
    QS_HIP(hipMalloc(&d_y_, rows * sizeof(float)));
    cap_ = rows;
  }
  QS_HIP(hipMemcpyAsync(d_x_, X.data(), rows * kIn * sizeof(float), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_forward, dim3(blocks_for(rows)), dim3(kThreads), 0, s, d_w_, d_x_, nullptr, (int)rows, nullptr,
                     nullptr, d_y_);
  QS_HIP(hipGetLastError());
  QS_HIP(hipMemcpyAsync(y.data(), d_y_, rows * sizeof(float), hipMemcpyDeviceToHost, s));
  QS_HIP(hipStreamSynchronize(s));
  return y;
}

}  // namespace mlp
}  // namespace qs
