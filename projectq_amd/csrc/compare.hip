// Element-wise comparison of two resident states (a measurement-harness extension: tests compare
// two simulators' states on one GPU without copying them to the host).
#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.hpp"

namespace qs {
namespace dev {
namespace {

template <typename Ra, typename Rb>
__global__ __launch_bounds__(kBlock) void k_compare(const amp_t<Ra> *__restrict__ a,
                                                    const amp_t<Rb> *__restrict__ b, uint64_t n,
                                                    double *partials) {
  __shared__ double lmax[kBlock / 64], lcnt[kBlock / 64];
  double dmax = 0.0, cnt = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const amp_t<Ra> x = a[i];
    const amp_t<Rb> y = b[i];
    const double dr = (double)x.x - (double)y.x, di = (double)x.y - (double)y.y;
    dmax = fmax(dmax, sqrt(dr * dr + di * di));
    // bitwise comparison of the stored values (meaningful when both have one precision)
    cnt += ((double)x.x != (double)y.x || (double)x.y != (double)y.y) ? 1.0 : 0.0;
  }
  for (int off = 32; off > 0; off >>= 1) {
    dmax = fmax(dmax, __shfl_down(dmax, off, 64));
    cnt += __shfl_down(cnt, off, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    lmax[wave] = dmax;
    lcnt[wave] = cnt;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w) {
      dmax = fmax(dmax, lmax[w]);
      cnt += lcnt[w];
    }
    partials[2 * blockIdx.x] = dmax;
    partials[2 * blockIdx.x + 1] = cnt;
  }
}

}  // namespace

void compare(const void *a, bool a_c64, const void *b, bool b_c64, uint64_t n, double *partials, hipStream_t s) {
  dim3 g(kReduceGrid), t(kBlock);
  if (!a_c64 && !b_c64)
    hipLaunchKernelGGL((k_compare<double, double>), g, t, 0, s, (const double2 *)a, (const double2 *)b, n, partials);
  else if (!a_c64 && b_c64)
    hipLaunchKernelGGL((k_compare<double, float>), g, t, 0, s, (const double2 *)a, (const float2 *)b, n, partials);
  else if (a_c64 && !b_c64)
    hipLaunchKernelGGL((k_compare<float, double>), g, t, 0, s, (const float2 *)a, (const double2 *)b, n, partials);
  else
    hipLaunchKernelGGL((k_compare<float, float>), g, t, 0, s, (const float2 *)a, (const float2 *)b, n, partials);
  QS_HIP(hipGetLastError());
}

}  // namespace dev
}  // namespace qs
