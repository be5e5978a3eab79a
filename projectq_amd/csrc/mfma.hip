// Dense k-qubit block pass on the matrix cores (complex64 states, k = 3, 4, 5).
//
// The VALU form of the pass (k_apply in kernels.hip) spends 2^k complex multiply-adds, 4 * 2^k
// FP32 FMAs, per amplitude.  Here the block is a real GEMM  Y = A X  (A: the 2D x 2D real form of
// the matrix, X: the groups' amplitudes as columns) on v_mfma_f32_32x32x2f32, which runs at twice
// the VALU FMA rate, with FP32 accumulation.  Operand layout and row orders: mfma_layout.hpp.
//
// Each wave handles 32 groups per loop turn.  Lane (n, h) loads the (re, im) pairs of half h of
// group n, so for consecutive groups a wave's load covers two 256-byte runs; ordinary (cached)
// loads and stores, since targets on low bits put the amplitudes of several groups in one line.
// The next turn's amplitudes are loaded before this turn's MFMA chain, so the matrix cores and
// the memory system work at the same time, and the matrix table is read from LDS at each step
// (k = 5: 16 KiB) rather than held in 64 VGPRs, which keeps five waves per SIMD.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "mfma_layout.hpp"

namespace qs {
namespace dev {
namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kWaves = kBlock / 64;

template <int K>
struct GroupOffsets {
  uint64_t o[1 << K];  // offset (in amplitudes) of amplitude j of a group from its base
};

__device__ __forceinline__ uint64_t insert_zeros_m(uint64_t g, const BitInsert &ins) {
  for (uint32_t i = 0; i < ins.count; ++i) {
    const uint32_t p = ins.pos[i];
    g = ((g >> p) << (p + 1)) | (g & ((uint64_t(1) << p) - 1));
  }
  return g;
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_apply_mfma(float2 *__restrict__ psi, uint64_t n_groups, BitInsert ins,
                                                       uint64_t cmask, GroupOffsets<K> off,
                                                       const float *__restrict__ table) {
  constexpr int D = 1 << K, H = D / 2, NB = mfma::blocks(K), S = mfma::steps(K);
  constexpr int QQ = K == 3 ? 2 : 4;  // register quads holding real output rows
  __shared__ float lds_a[NB * S * 64];
  for (int e = threadIdx.x; e < NB * S * 64; e += kBlock) lds_a[e] = table[e];
  __syncthreads();
  const int lane = threadIdx.x & 63, n = lane & 31, h = lane >> 5;
  const uint64_t half_off = h ? off.o[H] : 0;
  const uint64_t n_chunks = (n_groups + 31) / 32;
  const uint64_t first = (uint64_t)blockIdx.x * kWaves + (threadIdx.x >> 6), step = (uint64_t)gridDim.x * kWaves;
  auto group_ptr = [&](uint64_t c, bool *valid) {
    const uint64_t g = c * 32 + n;
    *valid = g < n_groups;
    return psi + ((insert_zeros_m(*valid ? g : 0, ins) | cmask) + half_off);
  };
  float2 x[H];
  bool valid = false;
  float2 *grp = psi;
  if (first < n_chunks) {
    grp = group_ptr(first, &valid);
#pragma unroll
    for (int m = 0; m < H; ++m) x[m] = valid ? grp[off.o[m]] : make_float2(0.f, 0.f);
  }
  for (uint64_t c = first; c < n_chunks; c += step) {
    asm volatile("" ::: "memory");  // the table is read from LDS inside the loop, not kept in VGPRs
    // the next turn's amplitudes: in flight during this turn's MFMA chain
    const uint64_t cn = c + step;
    bool vn = false;
    float2 *gn = psi;
    float2 xn[H];
    if (cn < n_chunks) {
      gn = group_ptr(cn, &vn);
#pragma unroll
      for (int m = 0; m < H; ++m) xn[m] = vn ? gn[off.o[m]] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      floatx16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int m = 0; m < H; ++m) {
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(lds_a[(b * S + 2 * m) * 64 + lane], x[m].x, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x2f32(lds_a[(b * S + 2 * m + 1) * 64 + lane], x[m].y, acc, 0, 0, 0);
      }
      if (valid) {
#pragma unroll
        for (int qq = 0; qq < QQ; ++qq)
#pragma unroll
          for (int s2 = 0; s2 < 2; ++s2)
            grp[off.o[8 * b + 2 * qq + s2]] = make_float2(acc[4 * qq + 2 * s2], acc[4 * qq + 2 * s2 + 1]);
      }
    }
#pragma unroll
    for (int m = 0; m < H; ++m) x[m] = xn[m];
    grp = gn;
    valid = vn;
  }
}

template <int K>
GroupOffsets<K> offsets_of(const uint32_t *tpos) {
  GroupOffsets<K> off;
  for (int j = 0; j < (1 << K); ++j) {
    uint64_t o = 0;
    for (int l = 0; l < K; ++l)
      if ((j >> l) & 1) o |= uint64_t(1) << tpos[l];
    off.o[j] = o;
  }
  return off;
}

template <int K>
void launch(float2 *psi, uint64_t n_groups, const BitInsert &ins, uint64_t cmask, const uint32_t *tpos,
            const float *d_table, hipStream_t s) {
  // one loop turn per wave would need n_chunks / 4 workgroups; past 4096 (16 waves on each of the
  // 256 CUs) the waves loop instead
  const uint64_t n_chunks = (n_groups + 31) / 32;
  uint64_t grid = (n_chunks + kWaves - 1) / kWaves;
  grid = grid < 1 ? 1 : (grid > 4096 ? 4096 : grid);
  hipLaunchKernelGGL((k_apply_mfma<K>), dim3((unsigned)grid), dim3(kBlock), 0, s, psi, n_groups, ins, cmask,
                     offsets_of<K>(tpos), d_table);
  QS_HIP(hipGetLastError());
}

}  // namespace

void apply_block_mfma(float2 *psi, uint64_t n_groups, const BitInsert &ins, uint64_t cmask, const uint32_t *tpos,
                      int k, const float *d_table, hipStream_t s) {
  switch (k) {
    case 3: launch<3>(psi, n_groups, ins, cmask, tpos, d_table, s); break;
    case 4: launch<4>(psi, n_groups, ins, cmask, tpos, d_table, s); break;
    case 5: launch<5>(psi, n_groups, ins, cmask, tpos, d_table, s); break;
    default: throw ValueError("apply_block_mfma: k must be 3, 4 or 5");
  }
}

}  // namespace dev
}  // namespace qs
