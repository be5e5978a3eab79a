// Pass-time cost model of the learned planner: a 3-layer MLP (24 -> 64 -> 64 -> 1, ReLU, fp32)
// trained and evaluated with the HIP kernels of mlp.hip.
//
// Weights (fp32, contiguous): W1[64][24], b1[64], W2[64][64], b2[64], w3[64], b3.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace qs {
namespace mlp {

constexpr int kIn = 24;
constexpr int kHidden = 64;
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kHidden * kIn;
constexpr int kW2 = kB1 + kHidden;
constexpr int kB2 = kW2 + kHidden * kHidden;
constexpr int kW3 = kB2 + kHidden;
constexpr int kB3 = kW3 + kHidden;
constexpr int kWeights = kB3 + 1;

// y[r] = f(X[r]) for rows r, on the GPU (host arrays in and out).
void predict(const float *w, const float *X, size_t rows, float *y);
// mean squared error over rows and its gradient w.r.t. every weight (host arrays).
void loss_grad(const float *w, const float *X, const float *T, size_t rows, float *loss, float *grad);
// Adam on minibatches of `batch` rows, rows reshuffled every epoch with mt19937(seed); He-normal
// initialisation from the same generator unless w_init is given.  Returns the final full-data
// loss; w_out receives the trained weights.
float train(const float *X, const float *T, size_t rows, int epochs, int batch, float lr, uint32_t seed,
            const float *w_init, float *w_out);

// A device-resident model for repeated inference from host feature rows.
class Model {
 public:
  explicit Model(const std::vector<float> &w);
  ~Model();
  Model(const Model &) = delete;
  Model &operator=(const Model &) = delete;
  std::vector<float> predict(const std::vector<float> &X, size_t rows, hipStream_t s);

 private:
  float *d_w_ = nullptr, *d_x_ = nullptr, *d_y_ = nullptr;
  size_t cap_ = 0;
};

}  // namespace mlp
}  // namespace qs
