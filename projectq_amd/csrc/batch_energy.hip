// Energies of a batch of circuits: <psi_c| H_c |psi_c> for every circuit c, H_c a sum of Pauli
// strings, FP64 sums whatever the state precision.
//
// Every amplitude i of a state meets every string t of its operator:
//   E = sum_i Re( conj(psi_i) sum_t c_t (-1)^{popcount(i & z_t)} psi_{i ^ x_t} )
// (the sign's constant part, popcount(x_t & z_t), is folded into c_t on the host).  A string with
// x_t != 0 pairs i with i ^ x_t, whose two contributions are complex conjugates (the string is
// Hermitian, its coefficient real): only the i with bit h_t, the highest bit of x_t, clear take
// part, with c_t doubled on the host -- half the partner loads and half the arithmetic.  The partner
// psi_{i ^ x_t} of 64 consecutive i is a permutation of one aligned 512-byte block, so each load
// is one contiguous block; the cost is the partner traffic from the caches (every string reads the
// whole state once), so:
//  * indices are 32-bit (batched circuits hold at most 30 qubits) and the per-string sign is the
//    parity of (i & z_t) applied as a sign-bit flip;
//  * eight partner loads are in flight before they are used;
//  * the host orders the workgroups so that all workgroups of one circuit run on one XCD
//    (workgroup w runs on XCD w mod 8): the partner reads of states up to 2 MiB then hit that
//    XCD's L2 instead of spreading every state over all eight.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace qs {
namespace dev {
namespace {

template <typename R>
__device__ __forceinline__ R flip_sign(R v, uint32_t odd);
template <>
__device__ __forceinline__ float flip_sign<float>(float v, uint32_t odd) {
  return __uint_as_float(__float_as_uint(v) ^ (odd << 31));
}
template <>
__device__ __forceinline__ double flip_sign<double>(double v, uint32_t odd) {
  return __longlong_as_double(__double_as_longlong(v) ^ ((unsigned long long)odd << 63));
}

// Neumaier-compensated accumulator
struct Compensated {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double v) {
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

__device__ __forceinline__ double workgroup_sum(double v, double *lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
  return t;
}

constexpr int kInFlight = 8;

template <typename R>
__global__ __launch_bounds__(kBlock) void k_batch_energy(const amp_t<R> *__restrict__ psi,
                                                         const EnergyJob *__restrict__ jobs,
                                                         const EnergyBlock *__restrict__ blocks,
                                                         const BatchTerm *__restrict__ terms,
                                                         double *__restrict__ partials) {
  __shared__ double lds[kBlock / 64];
  const EnergyBlock eb = blocks[blockIdx.x];
  if (eb.job == kNoJob) return;  // uniform over the workgroup
  const EnergyJob &job = jobs[eb.job];
  const amp_t<R> *__restrict__ st = psi + job.state_off;
  const BatchTerm *__restrict__ tm = terms + job.term_off;
  const uint32_t nt = job.n_terms;
  const uint32_t chunk = (uint32_t)job.chunk;
  const uint32_t lo = eb.chunk * chunk;
  const uint32_t hi = min((uint32_t)job.n_amps, lo + chunk);
  Compensated acc;
  for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) {
    const amp_t<R> a = st[i];
    double ur = 0.0, ui = 0.0;
    uint32_t t = 0;
    for (; t + kInFlight <= nt; t += kInFlight) {
      amp_t<R> b[kInFlight];
#pragma unroll
      for (int u = 0; u < kInFlight; ++u) {
        const BatchTerm &pt = tm[t + u];
        if (!(i & pt.h)) b[u] = st[i ^ pt.x];
      }
#pragma unroll
      for (int u = 0; u < kInFlight; ++u) {
        const BatchTerm &pt = tm[t + u];
        if (i & pt.h) continue;  // the pair's other member carries this contribution
        const uint32_t odd = __popc(i & pt.z) & 1u;
        const double br = (double)flip_sign<R>(b[u].x, odd), bi = (double)flip_sign<R>(b[u].y, odd);
        ur = fma(br, pt.c_re, fma(-bi, pt.c_im, ur));
        ui = fma(br, pt.c_im, fma(bi, pt.c_re, ui));
      }
    }
    for (; t < nt; ++t) {
      const BatchTerm &pt = tm[t];
      if (i & pt.h) continue;
      const amp_t<R> b = st[i ^ pt.x];
      const uint32_t odd = __popc(i & pt.z) & 1u;
      const double br = (double)flip_sign<R>(b.x, odd), bi = (double)flip_sign<R>(b.y, odd);
      ur = fma(br, pt.c_re, fma(-bi, pt.c_im, ur));
      ui = fma(br, pt.c_im, fma(bi, pt.c_re, ui));
    }
    acc.add(fma((double)a.x, ur, (double)a.y * ui));  // Re(conj(a) * u)
  }
  const double total = workgroup_sum(acc.value(), lds);
  if (threadIdx.x == 0) partials[job.first_block + eb.chunk] = total;
}

__global__ void k_batch_energy_finish(const EnergyJob *__restrict__ jobs, int n_jobs,
                                      const double *__restrict__ partials, double *energies) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_jobs) return;
  double s = 0.0;
  for (uint32_t b = 0; b < jobs[j].n_blocks; ++b) s += partials[jobs[j].first_block + b];
  energies[j] = s;
}

}  // namespace

template <typename R>
void batch_energy(const amp_t<R> *psi, const EnergyJob *jobs, const EnergyBlock *blocks, uint32_t n_blocks,
                  const BatchTerm *terms, double *partials, hipStream_t s) {
  if (n_blocks == 0) return;
  hipLaunchKernelGGL(k_batch_energy<R>, dim3(n_blocks), dim3(kBlock), 0, s, psi, jobs, blocks, terms, partials);
  QS_HIP(hipGetLastError());
}

void batch_energy_finish(const EnergyJob *jobs, int n_jobs, const double *partials, double *energies,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_batch_energy_finish, dim3((n_jobs + 255) / 256), dim3(256), 0, s, jobs, n_jobs, partials,
                     energies);
  QS_HIP(hipGetLastError());
}

template void batch_energy<float>(const amp_t<float> *, const EnergyJob *, const EnergyBlock *, uint32_t,
                                  const BatchTerm *, double *, hipStream_t);
template void batch_energy<double>(const amp_t<double> *, const EnergyJob *, const EnergyBlock *, uint32_t,
                                   const BatchTerm *, double *, hipStream_t);

}  // namespace dev
}  // namespace qs
