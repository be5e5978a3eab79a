// HIP kernels of the MI355X state-vector simulator (gfx950, wave64).
//
// Every kernel here is HBM-bound: one streaming pass over (a subspace of) the resident state.
// The dense-block kernel keeps the 2^k x 2^k matrix in scalar registers (uniform loads), holds
// the 2^k amplitudes of one group in VGPRs and streams each amplitude once in and once out.
// Reductions accumulate in FP64 whatever the state precision, with a fixed grid and a fixed
// final order so results are deterministic run to run.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "kernels.hpp"

namespace qs {
namespace dev {
namespace {

template <typename A>
__device__ __forceinline__ double re(const A &a) {
  return (double)a.x;
}
template <typename A>
__device__ __forceinline__ double im(const A &a) {
  return (double)a.y;
}
template <typename R>
__device__ __forceinline__ amp_t<R> make_amp(double x, double y) {
  amp_t<R> a;
  a.x = (R)x;
  a.y = (R)y;
  return a;
}

__device__ __forceinline__ uint64_t insert_zeros(uint64_t g, const BitInsert &ins) {
  for (uint32_t i = 0; i < ins.count; ++i) {
    uint32_t p = ins.pos[i];
    g = ((g >> p) << (p + 1)) | (g & ((uint64_t(1) << p) - 1));
  }
  return g;
}

__device__ __forceinline__ uint64_t insert_one_zero(uint64_t g, int p) {
  return ((g >> p) << (p + 1)) | (g & ((uint64_t(1) << p) - 1));
}

inline int grid_for(uint64_t work, int cap = 16384) {
  uint64_t g = (work + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > (uint64_t)cap) g = cap;
  return (int)g;
}

// Sum over a workgroup of kBlock threads; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
  __syncthreads();
  return t;
}

// Neumaier-compensated accumulator (used by the expectation value).
struct Kahan {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double v) {
    double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

// ------------------------------------------------------------------------------ gate blocks
template <int K>
struct Offsets {
  uint64_t o[1 << K];
};

template <typename R, int K>
__global__ __launch_bounds__(kBlock) void k_apply(amp_t<R> *__restrict__ psi, uint64_t n_groups,
                                                  BitInsert ins, uint64_t cmask, Offsets<K> off,
                                                  const amp_t<R> *__restrict__ m) {
  constexpr int D = 1 << K;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n_groups; g += stride) {
    const uint64_t base = insert_zeros(g, ins) | cmask;
    R vr[D], vi[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      amp_t<R> a = psi[base + off.o[j]];
      vr[j] = a.x;
      vi[j] = a.y;
    }
#pragma unroll(K <= 3 ? D : 1)
    for (int r = 0; r < D; ++r) {
      R ar = 0, ai = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const amp_t<R> mc = m[r * D + c];
        ar = fma(mc.x, vr[c], ar);
        ar = fma(-mc.y, vi[c], ar);
        ai = fma(mc.x, vi[c], ai);
        ai = fma(mc.y, vr[c], ai);
      }
      amp_t<R> o;
      o.x = ar;
      o.y = ai;
      psi[base + off.o[r]] = o;
    }
  }
}

template <typename R, int K>
void launch_apply(amp_t<R> *psi, uint64_t n_groups, const BitInsert &ins, uint64_t cmask,
                  const uint32_t *tpos, const amp_t<R> *m, hipStream_t s) {
  Offsets<K> off;
  for (int j = 0; j < (1 << K); ++j) {
    uint64_t o = 0;
    for (int l = 0; l < K; ++l)
      if ((j >> l) & 1) o |= uint64_t(1) << tpos[l];
    off.o[j] = o;
  }
  hipLaunchKernelGGL((k_apply<R, K>), dim3(grid_for(n_groups)), dim3(kBlock), 0, s, psi, n_groups,
                     ins, cmask, off, m);
  QS_HIP(hipGetLastError());
}

// ------------------------------------------------------------------------------ tile passes
template <typename R>
__device__ __forceinline__ uint32_t slot(uint32_t x);
template <>
__device__ __forceinline__ uint32_t slot<float>(uint32_t x) {
  return lds_slot_c64(x);
}
template <>
__device__ __forceinline__ uint32_t slot<double>(uint32_t x) {
  return lds_slot_c128(x);
}
// Both swizzles are linear over GF(2): for indices a, b without common bits,
// slot(a | b) = slot(a) ^ slot(b).  The kernels use this to address a group of amplitudes with
// one XOR per access: the per-thread part slot(lb) is computed once and the per-amplitude parts
// slot(offset) are uniform over the workgroup.  slot_bytes gives the byte offset in the tile.
template <typename R>
__device__ __forceinline__ uint32_t slot_bytes(uint32_t x) {
  return slot<R>(x) * (uint32_t)sizeof(amp_t<R>);
}
// Tiles are addressed by raw LDS byte address.  The tile kernels declare no static LDS, so the
// dynamic buffer (and the tile) starts at LDS address 0 and base ^ offset == base + offset; the
// XOR-composed offsets then need no add before the ds instruction.
template <typename T>
using lds_ptr = __attribute__((address_space(3))) T *;
__device__ __forceinline__ uint32_t lds_address(const void *p) {
  return (uint32_t)(uintptr_t)(lds_ptr<const char>)(p);
}
// T is a builtin vector type (16-byte tile streaming); amplitudes go through lds_*_amp.
template <typename T>
__device__ __forceinline__ T lds_load(uint32_t lds_byte_addr) {
  return *(lds_ptr<const T>)(lds_byte_addr);
}
template <typename T>
__device__ __forceinline__ void lds_store(uint32_t lds_byte_addr, const T &v) {
  *(lds_ptr<T>)(lds_byte_addr) = v;
}
template <typename R>
using re2_t = R __attribute__((ext_vector_type(2)));
template <typename R>
__device__ __forceinline__ amp_t<R> lds_load_amp(uint32_t lds_byte_addr) {
  const re2_t<R> v = lds_load<re2_t<R>>(lds_byte_addr);
  amp_t<R> a;
  a.x = v.x;
  a.y = v.y;
  return a;
}
template <typename R>
__device__ __forceinline__ void lds_store_amp(uint32_t lds_byte_addr, const amp_t<R> &a) {
  re2_t<R> v;
  v.x = a.x;
  v.y = a.y;
  lds_store<re2_t<R>>(lds_byte_addr, v);
}

// One TileOp of any width straight on the LDS tile (NT threads per workgroup).
template <typename R, int K, int NT = kBlock>
__device__ __forceinline__ void tile_apply(amp_t<R> *__restrict__ tile, uint32_t t,
                                           const TileOp *__restrict__ op,
                                           const amp_t<R> *__restrict__ m) {
  constexpr int D = 1 << K;
  const uint32_t nfixed = op->nfixed, lcmask = op->lcmask;
  const uint32_t ngroups = 1u << (t - nfixed);
  for (uint32_t g = threadIdx.x; g < ngroups; g += NT) {
    uint32_t lb = g;
    for (uint32_t i = 0; i < nfixed; ++i) {
      const uint32_t p = op->fixed[i];
      lb = ((lb >> p) << (p + 1)) | (lb & ((1u << p) - 1));
    }
    lb |= lcmask;
    uint32_t sb = lds_address(tile) + slot_bytes<R>(lb);
    asm volatile("" : "+v"(sb));
    R vr[D], vi[D];
#pragma unroll
    for (int j = 0; j < D; ++j) {
      const amp_t<R> a = lds_load_amp<R>(sb ^ slot_bytes<R>(op->off[j]));
      vr[j] = a.x;
      vi[j] = a.y;
    }
    asm volatile("" : "+v"(sb));
#pragma unroll(K <= 3 ? D : 1)
    for (int r = 0; r < D; ++r) {
      R ar = 0, ai = 0;
#pragma unroll
      for (int c = 0; c < D; ++c) {
        const amp_t<R> mc = m[r * D + c];
        ar = fma(mc.x, vr[c], ar);
        ar = fma(-mc.y, vi[c], ar);
        ai = fma(mc.x, vi[c], ai);
        ai = fma(mc.y, vr[c], ai);
      }
      amp_t<R> o;
      o.x = ar;
      o.y = ai;
      lds_store_amp<R>(sb ^ slot_bytes<R>(op->off[r]), o);
    }
  }
}

// ---- register-stage ops: NR = 2^RB amplitudes of one thread, register index bit i <-> stage
// bit i.  C is the op's control among the register bits (-1: none); it is a compile-time
// constant, so a controlled op touches only the indices with bit C set and a controlled X is a
// renaming of registers.
//
// Controls on tile bits outside the register bits hold for some threads and not others.  They do
// not branch: ops with such controls run the PRED form, in which every lane runs the op and lanes
// whose controls fail (`act` false) apply the identity -- the matrix is blended with the
// identity, or the results are selected.  With no divergent branch in the op loop the
// op dispatch stays uniform control flow, which the compiler lays out without copying the
// amplitude registers at every merge (with -structurizecfg-skip-uniform-regions).
__host__ __device__ constexpr bool reg_ctrl_ok(int i, int C) { return C < 0 || ((i >> (C < 0 ? 0 : C)) & 1); }

template <typename R>
__device__ __forceinline__ R blend(bool act, R v, R id) { return act ? v : id; }

template <typename R, int RB, bool PRED, int C, int B>
__device__ __forceinline__ void rop_general(R *vr, R *vi, const amp_t<R> *__restrict__ m, bool act) {
  R ar = m[0].x, ai = m[0].y, br = m[1].x, bi = m[1].y;
  R cr = m[2].x, ci = m[2].y, dr = m[3].x, di = m[3].y;
  if constexpr (PRED) {
    ar = blend<R>(act, ar, 1), ai = blend<R>(act, ai, 0), br = blend<R>(act, br, 0), bi = blend<R>(act, bi, 0);
    cr = blend<R>(act, cr, 0), ci = blend<R>(act, ci, 0), dr = blend<R>(act, dr, 1), di = blend<R>(act, di, 0);
  }
#pragma unroll
  for (int i = 0; i < (1 << RB); ++i) {
    if ((i & (1 << B)) || !reg_ctrl_ok(i, C)) continue;
    const int j = i | (1 << B);
    const R xr = vr[i], xi = vi[i], yr = vr[j], yi = vi[j];
    vr[i] = fma(ar, xr, fma(-ai, xi, fma(br, yr, -bi * yi)));
    vi[i] = fma(ar, xi, fma(ai, xr, fma(br, yi, bi * yr)));
    vr[j] = fma(cr, xr, fma(-ci, xi, fma(dr, yr, -di * yi)));
    vi[j] = fma(cr, xi, fma(ci, xr, fma(dr, yi, di * yr)));
  }
}
template <typename R, int RB, bool PRED, int C, int B>
__device__ __forceinline__ void rop_real(R *vr, R *vi, const amp_t<R> *__restrict__ m, bool act) {
  R a = m[0].x, b = m[1].x, c = m[2].x, d = m[3].x;
  if constexpr (PRED) a = blend<R>(act, a, 1), b = blend<R>(act, b, 0), c = blend<R>(act, c, 0), d = blend<R>(act, d, 1);
#pragma unroll
  for (int i = 0; i < (1 << RB); ++i) {
    if ((i & (1 << B)) || !reg_ctrl_ok(i, C)) continue;
    const int j = i | (1 << B);
    const R xr = vr[i], xi = vi[i], yr = vr[j], yi = vi[j];
    vr[i] = fma(a, xr, b * yr);
    vi[i] = fma(a, xi, b * yi);
    vr[j] = fma(c, xr, d * yr);
    vi[j] = fma(c, xi, d * yi);
  }
}
template <typename R, int RB, bool PRED, int C, int B>
__device__ __forceinline__ void rop_diag(R *vr, R *vi, const amp_t<R> *__restrict__ m, bool act) {
  R ar = m[0].x, ai = m[0].y, dr = m[3].x, di = m[3].y;
  if constexpr (PRED) ar = blend<R>(act, ar, 1), ai = blend<R>(act, ai, 0), dr = blend<R>(act, dr, 1), di = blend<R>(act, di, 0);
#pragma unroll
  for (int i = 0; i < (1 << RB); ++i) {
    if (!reg_ctrl_ok(i, C)) continue;
    const R fr = (i & (1 << B)) ? dr : ar, fi = (i & (1 << B)) ? di : ai;
    const R xr = vr[i], xi = vi[i];
    vr[i] = fma(fr, xr, -fi * xi);
    vi[i] = fma(fr, xi, fi * xr);
  }
}
template <typename R, int RB, bool PRED, int C, int B>
__device__ __forceinline__ void rop_anti(R *vr, R *vi, const amp_t<R> *__restrict__ m, bool act) {
  const R br = m[1].x, bi = m[1].y, cr = m[2].x, ci = m[2].y;
#pragma unroll
  for (int i = 0; i < (1 << RB); ++i) {
    if ((i & (1 << B)) || !reg_ctrl_ok(i, C)) continue;
    const int j = i | (1 << B);
    const R xr = vr[i], xi = vi[i], yr = vr[j], yi = vi[j];
    const R nr_i = fma(br, yr, -bi * yi), ni_i = fma(br, yi, bi * yr);
    const R nr_j = fma(cr, xr, -ci * xi), ni_j = fma(cr, xi, ci * xr);
    vr[i] = PRED ? blend<R>(act, nr_i, xr) : nr_i;
    vi[i] = PRED ? blend<R>(act, ni_i, xi) : ni_i;
    vr[j] = PRED ? blend<R>(act, nr_j, yr) : nr_j;
    vi[j] = PRED ? blend<R>(act, ni_j, yi) : ni_j;
  }
}
template <typename R, int RB, bool PRED, int C, int B>
__device__ __forceinline__ void rop_x(R *vr, R *vi, const amp_t<R> *__restrict__, bool act) {
  if constexpr (PRED) {
#pragma unroll
    for (int i = 0; i < (1 << RB); ++i) {
      if ((i & (1 << B)) || !reg_ctrl_ok(i, C)) continue;
      const int j = i | (1 << B);
      const R xr = vr[i], xi = vi[i], yr = vr[j], yi = vi[j];
      vr[i] = blend<R>(act, yr, xr);
      vi[i] = blend<R>(act, yi, xi);
      vr[j] = blend<R>(act, xr, yr);
      vi[j] = blend<R>(act, xi, yi);
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < (1 << RB); ++i) {
    if ((i & (1 << B)) || !reg_ctrl_ok(i, C)) continue;
    const int j = i | (1 << B);
    const R xr = vr[i], xi = vi[i];
    vr[i] = vr[j];
    vi[i] = vi[j];
    vr[j] = xr;
    vi[j] = xi;
  }
}
// Z (diag(1, -1)): the amplitudes with bit B set change sign, in place (an XOR of the sign bit;
// predicated lanes XOR with zero).  A CNOT between 1-qubit gates on its target reaches the tile
// kernels as H, controlled Z, H with the H's merged into those gates (cnot_to_cz in planner.cpp).
template <typename R>
__device__ __forceinline__ R flip_sign(R x, bool act) {
  if constexpr (sizeof(R) == 4) return __uint_as_float(__float_as_uint(x) ^ (act ? 0x80000000u : 0u));
  else return __longlong_as_double(__double_as_longlong(x) ^ (act ? (1ll << 63) : 0ll));
}
template <typename R, int RB, bool PRED, int C, int B>
__device__ __forceinline__ void rop_z(R *vr, R *vi, const amp_t<R> *__restrict__, bool act) {
#pragma unroll
  for (int i = 0; i < (1 << RB); ++i) {
    if (!(i & (1 << B)) || !reg_ctrl_ok(i, C)) continue;
    vr[i] = flip_sign<R>(vr[i], PRED ? act : true);
    vi[i] = flip_sign<R>(vi[i], PRED ? act : true);
  }
}
template <typename R, int RB, bool PRED, int C, int B0, int B1>
__device__ __forceinline__ void rop_general2(R *vr, R *vi, const amp_t<R> *__restrict__ m, bool act) {
  R mr[16], mi[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    mr[k] = m[k].x;
    mi[k] = m[k].y;
    if constexpr (PRED) mr[k] = blend<R>(act, mr[k], (k % 5) == 0 ? R(1) : R(0)), mi[k] = blend<R>(act, mi[k], 0);
  }
#pragma unroll
  for (int i = 0; i < (1 << RB); ++i) {
    if ((i & ((1 << B0) | (1 << B1))) || !reg_ctrl_ok(i, C)) continue;
    const int idx[4] = {i, i | (1 << B0), i | (1 << B1), i | (1 << B0) | (1 << B1)};
    R xr[4], xi[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      xr[c] = vr[idx[c]];
      xi[c] = vi[idx[c]];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      R sr = 0, si = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        sr = fma(mr[r * 4 + c], xr[c], fma(-mi[r * 4 + c], xi[c], sr));
        si = fma(mr[r * 4 + c], xi[c], fma(mi[r * 4 + c], xr[c], si));
      }
      vr[idx[r]] = sr;
      vi[idx[r]] = si;
    }
  }
}
template <typename R, int RB, bool PRED, int C, int B0, int B1>
__device__ __forceinline__ void rop_diag2(R *vr, R *vi, const amp_t<R> *__restrict__ m, bool act) {
  R fr[4], fi[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    fr[s] = m[s * 5].x;  // diagonal element s of the 4x4 matrix
    fi[s] = m[s * 5].y;
    if constexpr (PRED) fr[s] = blend<R>(act, fr[s], 1), fi[s] = blend<R>(act, fi[s], 0);
  }
#pragma unroll
  for (int i = 0; i < (1 << RB); ++i) {
    if (!reg_ctrl_ok(i, C)) continue;
    const int s = ((i >> B0) & 1) | (((i >> B1) & 1) << 1);
    const R xr = vr[i], xi = vi[i];
    vr[i] = fma(fr[s], xr, -fi[s] * xi);
    vi[i] = fma(fr[s], xi, fi[s] * xr);
  }
}

// Dispatch on the kind, register bits and register control of an op (cb = 1 + control bit, 0 =
// none), in one function so that the amplitude registers stay in registers.  Combinations the
// planner never emits (a control on a target, bits >= RB) do nothing.
#define QS_R1(FN, C, B)                                          \
  if constexpr ((B) < RB && (C) < RB && (C) != (B)) FN<R, RB, PRED, C, B>(vr, vi, m, act);
#define QS_R1C(FN, B)                                            \
  switch (cb) {                                                  \
    case 0: QS_R1(FN, -1, B) break;                              \
    case 1: QS_R1(FN, 0, B) break;                               \
    case 2: QS_R1(FN, 1, B) break;                               \
    case 3: QS_R1(FN, 2, B) break;                               \
    case 4: QS_R1(FN, 3, B) break;                               \
    default: QS_R1(FN, 4, B) break;                              \
  }
#define QS_R1B(FN)                                               \
  switch (b0) {                                                  \
    case 0: QS_R1C(FN, 0) break;                                 \
    case 1: QS_R1C(FN, 1) break;                                 \
    case 2: QS_R1C(FN, 2) break;                                 \
    case 3: QS_R1C(FN, 3) break;                                 \
    default: QS_R1C(FN, 4) break;                                \
  }
#define QS_R2(FN, C, B0, B1)                                                            \
  if constexpr ((B1) < RB && (C) < RB && (C) != (B0) && (C) != (B1)) FN<R, RB, PRED, C, B0, B1>(vr, vi, m, act);
#define QS_R2C(FN, B0, B1)                                       \
  switch (cb) {                                                  \
    case 0: QS_R2(FN, -1, B0, B1) break;                         \
    case 1: QS_R2(FN, 0, B0, B1) break;                          \
    case 2: QS_R2(FN, 1, B0, B1) break;                          \
    case 3: QS_R2(FN, 2, B0, B1) break;                          \
    case 4: QS_R2(FN, 3, B0, B1) break;                          \
    default: QS_R2(FN, 4, B0, B1) break;                         \
  }
#define QS_R2B(FN)                                               \
  switch (b0 * 8 + b1) {                                         \
    case 0 * 8 + 1: QS_R2C(FN, 0, 1) break;                      \
    case 0 * 8 + 2: QS_R2C(FN, 0, 2) break;                      \
    case 0 * 8 + 3: QS_R2C(FN, 0, 3) break;                      \
    case 1 * 8 + 2: QS_R2C(FN, 1, 2) break;                      \
    case 1 * 8 + 3: QS_R2C(FN, 1, 3) break;                      \
    case 2 * 8 + 3: QS_R2C(FN, 2, 3) break;                      \
    case 0 * 8 + 4: QS_R2C(FN, 0, 4) break;                      \
    case 1 * 8 + 4: QS_R2C(FN, 1, 4) break;                      \
    case 2 * 8 + 4: QS_R2C(FN, 2, 4) break;                      \
    default: QS_R2C(FN, 3, 4) break;                             \
  }

// TWO = false: a kernel for passes with no 2-qubit register op (the host picks it per pass); the
// dispatch then holds only the 1-qubit forms, and the register allocator works on far fewer arms.
template <typename R, int RB, bool PRED, bool TWO = true>
__device__ __forceinline__ void reg_op(uint32_t code, R *vr, R *vi, const amp_t<R> *__restrict__ m, bool act) {
  uint32_t kind = (code >> 8) & 15u;
  const uint32_t b0 = code & 15u, b1 = (code >> 4) & 15u, cb = code >> 12;
  (void)b1;
  // the full kernel runs a Z op as the diagonal op it is (its matrix is diag(1, -1)): a set of Z
  // arms beside the 2-qubit ones measured 4% slower on every pass
  if constexpr (TWO) kind = kind == kOp1Z ? kOp1Diag : kind;
  switch (kind) {
    case kOp1General: QS_R1B(rop_general) break;
    case kOp1Real: QS_R1B(rop_real) break;
    case kOp1Diag: QS_R1B(rop_diag) break;
    case kOp1Anti: QS_R1B(rop_anti) break;
    case kOp1X: QS_R1B(rop_x) break;
    case kOp1Z:
      if constexpr (!TWO) { QS_R1B(rop_z) }
      break;
    default:
      if constexpr (TWO) {
        if (kind == kOp2General) QS_R2B(rop_general2)
        else QS_R2B(rop_diag2)
      }
      break;
  }
}
#undef QS_R1
#undef QS_R1C
#undef QS_R1B
#undef QS_R2
#undef QS_R2C
#undef QS_R2B

// One register op: global controls are uniform over the block (a branch); tile-bit controls are
// per lane (the predicated form).
template <typename R, int RB, bool TWO = true>
__device__ __forceinline__ void reg_apply(const RegOp *__restrict__ op, uint64_t base, uint32_t lb, R *vr, R *vi,
                                          const amp_t<R> *__restrict__ mats) {
  const uint64_t gc = op->gcmask;
  const uint32_t lc = op->lcmask;
  if ((base & gc) == gc) {
    if (lc == 0) reg_op<R, RB, false, TWO>(op->code, vr, vi, mats + op->mat, true);
    else reg_op<R, RB, true, TWO>(op->code, vr, vi, mats + op->mat, (lb & lc) == lc);
  }
}

// TURN ops per turn of the op loop.  With 2 the register allocator moves the amplitudes between
// register sets once per turn rather than once per op (worth about 2% on the pass kernel), at
// the price of a second copy of the op dispatch; kernels off the hot path use 1.
template <typename R, int RB, int TURN, int NT = kBlock, bool TWO = true>
__device__ __forceinline__ void reg_stage(amp_t<R> *__restrict__ tile, uint32_t t, uint64_t base,
                                          const TileStage *__restrict__ st,
                                          const RegOp *__restrict__ rops,
                                          const amp_t<R> *__restrict__ mats) {
  constexpr int NR = 1 << RB;
  uint32_t soff[NR];  // byte offset (swizzled) of register i relative to the group's base
#pragma unroll
  for (int i = 0; i < NR; ++i) soff[i] = st->soff[i];
  const uint32_t ngroups = 1u << (t - RB);
  const uint32_t ob = st->begin, oe = st->end;
  for (uint32_t g = threadIdx.x; g < ngroups; g += NT) {
    uint32_t lb = g;
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      const uint32_t p = st->fixed[b];
      lb = ((lb >> p) << (p + 1)) | (lb & ((1u << p) - 1));
    }
    uint32_t sb = lds_address(tile) + slot_bytes<R>(lb);
    asm volatile("" : "+v"(sb));  // keeps the address math at one XOR per access
    R vr[NR], vi[NR];
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      const amp_t<R> a = lds_load_amp<R>(sb ^ soff[i]);
      vr[i] = a.x;
      vi[i] = a.y;
    }
    for (uint32_t o = ob; o < oe; o += TURN) {
      reg_apply<R, RB, TWO>(rops + o, base, lb, vr, vi, mats);
      if constexpr (TURN == 2) {
        if (o + 1 < oe) reg_apply<R, RB, TWO>(rops + o + 1, base, lb, vr, vi, mats);
      }
    }
    // Recompute the LDS addresses for the write-back instead of keeping 2^RB of them live across
    // the ops (the empty asm makes sb opaque to common-subexpression elimination).
    asm volatile("" : "+v"(sb));
#pragma unroll
    for (int i = 0; i < NR; ++i) {
      amp_t<R> a;
      a.x = vr[i];
      a.y = vi[i];
      lds_store_amp<R>(sb ^ soff[i], a);
    }
  }
}

// 16-byte vectors of amplitudes used to stream tiles (2 x complex64 or 1 x complex128)
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
template <typename R>
struct Vec16 {
  using type = f32x4;
  static constexpr int n = 2;
};
template <>
struct Vec16<double> {
  using type = f64x2;
  static constexpr int n = 1;
};

// The LDS side of tile streaming: one 16-byte global vector holds one complex128 amplitude or
// the complex64 pair (l, l + 1).  The swizzle keeps such a pair adjacent but may swap it
// (slot(l + 1) = slot(l) ^ 1), so complex64 pairs move as two 8-byte accesses.
template <typename R>
__device__ __forceinline__ void tile_put16(uint32_t addr, const typename Vec16<R>::type &v) {
  if constexpr (sizeof(R) == 4) {
    re2_t<float> a, b;
    a.x = v.x;
    a.y = v.y;
    b.x = v.z;
    b.y = v.w;
    lds_store<re2_t<float>>(addr, a);
    lds_store<re2_t<float>>(addr ^ 8u, b);
  } else {
    lds_store<typename Vec16<R>::type>(addr, v);
  }
}
// The same read when the stored tile relabels qubits: the pair (l, l + 1) of the stored tile comes
// from computed-tile slots addr and addr ^ pair (pair = the slot bytes of the partner's bit).
template <typename R>
__device__ __forceinline__ typename Vec16<R>::type tile_get16p(uint32_t addr, uint32_t pair) {
  if constexpr (sizeof(R) == 4) {
    const re2_t<float> a = lds_load<re2_t<float>>(addr), b = lds_load<re2_t<float>>(addr ^ pair);
    typename Vec16<R>::type v;
    v.x = a.x;
    v.y = a.y;
    v.z = b.x;
    v.w = b.y;
    return v;
  } else {
    return lds_load<typename Vec16<R>::type>(addr);
  }
}
template <typename R>
__device__ __forceinline__ typename Vec16<R>::type tile_get16(uint32_t addr) {
  if constexpr (sizeof(R) == 4) {
    const re2_t<float> a = lds_load<re2_t<float>>(addr), b = lds_load<re2_t<float>>(addr ^ 8u);
    typename Vec16<R>::type v;
    v.x = a.x;
    v.y = a.y;
    v.z = b.x;
    v.w = b.y;
    return v;
  } else {
    return lds_load<typename Vec16<R>::type>(addr);
  }
}

// Streaming accesses of the tile passes: every amplitude is read and written exactly once per
// pass, so the accesses carry the non-temporal hint unless QS_TILE_TEMPORAL is defined.
template <typename V>
__device__ __forceinline__ V stream_load(const V *p) {
#ifdef QS_TILE_TEMPORAL
  return *p;
#else
  return __builtin_nontemporal_load(p);
#endif
}
template <typename V>
__device__ __forceinline__ void stream_store(V *p, const V &v) {
#ifdef QS_TILE_TEMPORAL
  *p = v;
#else
  __builtin_nontemporal_store(v, p);
#endif
}

// Global index of the first amplitude of tile `tix`: the low L bits and the high tile positions
// (ascending, all >= L) are zero and the remaining bits hold tix.  The loop has a fixed trip count
// so the positions are read at static offsets of the kernel argument (scalar loads).
__device__ __forceinline__ uint64_t tile_base(uint64_t tix, const TilePass &P) {
  uint64_t g = tix << P.L;
#pragma unroll
  for (int b = 0; b < kTileMaxHigh; ++b)
    if (b < (int)P.h) g = insert_one_zero(g, (int)P.hpos[b]);
  return g;
}

// 16-byte accesses each thread keeps in flight while streaming a tile.
constexpr int kStreamChunk = 8;

// Tile-local index of the computed tile that index x of the stored tile takes (the relabelling of
// P.sperm; x itself when there is none).  GF(2)-linear in x.
__device__ __forceinline__ uint32_t store_source(uint32_t x, const TilePass &P) {
  uint32_t y = 0;
#pragma unroll
  for (int b = 0; b < kTileMaxBits; ++b)
    if (b < (int)P.t) y |= ((x >> b) & 1u) << P.sperm[b];
  return y;
}

// The stages of a pass on one tile resident in LDS, a workgroup barrier after each.
template <typename R, int RB, int TURN, int NT = kBlock, bool TWO = true>
__device__ __forceinline__ void run_stages(amp_t<R> *__restrict__ tile, uint32_t t, uint64_t base, uint32_t n_stages,
                                           const TileStage *__restrict__ stages, const RegOp *__restrict__ rops,
                                           const TileOp *__restrict__ tops, const amp_t<R> *__restrict__ mats) {
  for (uint32_t s = 0; s < n_stages; ++s) {
    const TileStage *st = stages + s;
    if (st->kind == kStageRegisters) {
      reg_stage<R, RB, TURN, NT, TWO>(tile, t, base, st, rops, mats);
    } else {
      const TileOp *op = tops + st->begin;
      const uint64_t gc = op->gcmask;
      if ((base & gc) == gc) {
        const amp_t<R> *m = mats + op->mat;
        switch (op->k) {
          case 1: tile_apply<R, 1, NT>(tile, t, op, m); break;
          case 2: tile_apply<R, 2, NT>(tile, t, op, m); break;
          case 3: tile_apply<R, 3, NT>(tile, t, op, m); break;
          case 4: tile_apply<R, 4, NT>(tile, t, op, m); break;
          default: tile_apply<R, 5, NT>(tile, t, op, m); break;
        }
      }
    }
    __syncthreads();
  }
}

// One tile of one pass: load the 2^t amplitudes into LDS, run the stages, store them back.
// `seg` holds the 2^h segment offsets of the pass (shared memory, filled by the caller).
// FULL: 2^t is a multiple of kStreamChunk * kBlock * VN, so every streamed vector exists and the
// accesses need no guard (straight-line code lets the memory counters be waited on exactly).
template <typename R, int RB, bool FULL, bool TWO = true>
__device__ __forceinline__ void tile_body(amp_t<R> *__restrict__ psi, uint64_t tix, const TilePass &P,
                                          const uint64_t *seg, amp_t<R> *__restrict__ tile,
                                          const TileStage *__restrict__ stages,
                                          const RegOp *__restrict__ rops, const TileOp *__restrict__ tops,
                                          const amp_t<R> *__restrict__ mats) {
  using V = typename Vec16<R>::type;
  constexpr int VN = Vec16<R>::n;
  constexpr uint32_t kStep = kBlock * VN;
  const uint32_t t = P.t, L = P.L;
  const uint32_t tsize = 1u << t, segmask = (1u << L) - 1;
  const uint64_t base = tile_base(tix, P);
  for (uint32_t l0 = threadIdx.x * VN; l0 < tsize; l0 += kStreamChunk * kStep) {
    const uint32_t s0 = lds_address(tile) + slot_bytes<R>(l0);  // l0 shares no bit with j * kStep
    V v[kStreamChunk];
#pragma unroll
    for (int j = 0; j < kStreamChunk; ++j) {
      const uint32_t l = l0 + j * kStep;
      if (FULL || l < tsize) v[j] = stream_load(reinterpret_cast<const V *>(psi + (base | seg[l >> L] | (l & segmask))));
    }
#pragma unroll
    for (int j = 0; j < kStreamChunk; ++j) {
      const uint32_t l = l0 + j * kStep;
      if (FULL || l < tsize) tile_put16<R>(s0 ^ slot_bytes<R>(j * kStep), v[j]);
    }
  }
  __syncthreads();
  run_stages<R, RB, 1, kBlock, TWO>(tile, t, base, P.n_stages, stages, rops, tops, mats);
  // the store reads the computed tile through the relabelling (linear: sources of disjoint bit
  // sets combine by XOR)
  const uint32_t src_thread = store_source(threadIdx.x * VN, P), pair = slot_bytes<R>(store_source(1, P));
  for (uint32_t l0 = threadIdx.x * VN; l0 < tsize; l0 += kStreamChunk * kStep) {
    const uint32_t s0 = lds_address(tile) + slot_bytes<R>(src_thread ^ store_source(l0 - threadIdx.x * VN, P));
    V v[kStreamChunk];
#pragma unroll
    for (int j = 0; j < kStreamChunk; ++j) {
      const uint32_t l = l0 + j * kStep;
      if (FULL || l < tsize) v[j] = tile_get16p<R>(s0 ^ slot_bytes<R>(store_source(j * kStep, P)), pair);
    }
#pragma unroll
    for (int j = 0; j < kStreamChunk; ++j) {
      const uint32_t l = l0 + j * kStep;
      if (FULL || l < tsize) stream_store(reinterpret_cast<V *>(psi + (base | seg[l >> L] | (l & segmask))), v[j]);
    }
  }
  __syncthreads();
}

template <int NT = kBlock>
__device__ __forceinline__ void fill_segments(const TilePass &P, uint64_t *seg) {
  for (uint32_t sgi = threadIdx.x; sgi < (1u << P.h); sgi += NT) {
    uint64_t o = 0;
    for (uint32_t b = 0; b < P.h; ++b)
      if ((sgi >> b) & 1) o |= uint64_t(1) << P.hpos[b];
    seg[sgi] = o;
  }
  __syncthreads();
}

// TWO = false: the instantiation for passes without 2-qubit register ops (see reg_op)
template <typename R, int RB, bool FULL, bool TWO = true>
__global__ __launch_bounds__(kBlock) void k_tile_pass(amp_t<R> *__restrict__ psi, uint64_t n_tiles,
                                                      TilePass P, const TileStage *__restrict__ stages,
                                                      const RegOp *__restrict__ rops,
                                                      const TileOp *__restrict__ tops,
                                                      const amp_t<R> *__restrict__ mats) {
  extern __shared__ __align__(16) unsigned char smem[];
  // the tile starts at LDS address 0 (no static LDS), so tile offsets need no base add;
  // the 2^h segment offsets follow the tile
  amp_t<R> *tile = reinterpret_cast<amp_t<R> *>(smem);
  uint64_t *seg = reinterpret_cast<uint64_t *>(smem + (sizeof(amp_t<R>) << P.t));
  fill_segments(P, seg);
  for (uint64_t tix = blockIdx.x; tix < n_tiles; tix += gridDim.x)
    tile_body<R, RB, FULL, TWO>(psi, tix, P, seg, tile, stages, rops, tops, mats);
}

// Software-pipelined tile pass: while the stages of tile i run out of LDS, the loads of the
// workgroup's next tile are already in flight into registers, so HBM traffic overlaps the in-tile
// arithmetic.  Used when a thread's share of a tile is exactly kStreamChunk 16-byte vectors
// (2^t = kStreamChunk * NT * VN), so all streaming is unguarded straight-line code.  NT = 256
// serves 2^12-amplitude complex64 tiles (32 KiB of LDS, four workgroups per CU), NT = 512 tiles of
// 2^13 (64 KiB, two workgroups per CU): the same per-thread work and the same waves per CU.
template <typename R, int RB, int NT, bool TWO>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void k_tile_pass_pf(amp_t<R> *__restrict__ psi, uint64_t n_tiles,
                                                         TilePass P, const TileStage *__restrict__ stages,
                                                         const RegOp *__restrict__ rops,
                                                         const TileOp *__restrict__ tops,
                                                         const amp_t<R> *__restrict__ mats) {
  using V = typename Vec16<R>::type;
  constexpr int VN = Vec16<R>::n;
  constexpr uint32_t kStep = NT * VN;
  extern __shared__ __align__(16) unsigned char smem[];
  // the tile starts at LDS address 0 (no static LDS), so tile offsets need no base add;
  // the 2^h segment offsets follow the tile
  amp_t<R> *tile = reinterpret_cast<amp_t<R> *>(smem);
  uint64_t *seg = reinterpret_cast<uint64_t *>(smem + (sizeof(amp_t<R>) << P.t));
  fill_segments<NT>(P, seg);
  uint64_t tix = blockIdx.x;
  if (tix >= n_tiles) return;  // uniform over the workgroup
  const uint32_t t = P.t, L = P.L;
  const uint32_t segmask = (1u << L) - 1;
  const uint32_t l0 = threadIdx.x * VN;
  const uint32_t s0 = lds_address(tile) + slot_bytes<R>(l0);  // l0 shares no bit with j * kStep
  // the store reads the computed tile through the relabelling of P.sperm (identity: s0, j * kStep)
  const uint32_t ss0 = lds_address(tile) + slot_bytes<R>(store_source(l0, P));
  const uint32_t spair = slot_bytes<R>(store_source(1, P));
  // j * kStep for j < 8 spans three bits: its source is the XOR of theirs
  const uint32_t sbit0 = slot_bytes<R>(store_source(kStep, P)), sbit1 = slot_bytes<R>(store_source(2 * kStep, P)),
                 sbit2 = slot_bytes<R>(store_source(4 * kStep, P));
  static_assert(kStreamChunk == 8, "three step bits");
  // Each turn of the loop runs the stages of the resident tile and stores it, then moves the
  // registers holding the next tile (loaded during those stages) into LDS and starts loading the
  // tile after.  At that wait only the older loads must be complete, not the stores just issued.
  uint64_t base = tile_base(tix, P);
  V v[kStreamChunk];
#pragma unroll
  for (int j = 0; j < kStreamChunk; ++j) {
    const uint32_t l = l0 + j * kStep;
    v[j] = stream_load(reinterpret_cast<const V *>(psi + (base | seg[l >> L] | (l & segmask))));
  }
#pragma unroll
  for (int j = 0; j < kStreamChunk; ++j) tile_put16<R>(s0 ^ slot_bytes<R>(j * kStep), v[j]);
  __syncthreads();
  uint64_t next = tix + gridDim.x;
  uint64_t nbase = tile_base(next < n_tiles ? next : tix, P);
  auto prefetch = [&](uint64_t from) {
    // the segment offsets are re-read from LDS for every tile rather than kept live across the
    // stages (the empty asm keeps the compiler from hoisting them into registers)
    uint32_t lq = l0;
    asm volatile("" : "+v"(lq));
#pragma unroll
    for (int j = 0; j < kStreamChunk; ++j) {
      const uint32_t l = lq + j * kStep;
      v[j] = stream_load(reinterpret_cast<const V *>(psi + (from | seg[l >> L] | (l & segmask))));
    }
  };
  if (next < n_tiles) prefetch(nbase);
  while (true) {
    run_stages<R, RB, 2, NT, TWO>(tile, t, base, P.n_stages, stages, rops, tops, mats);
    uint32_t ls = l0;
    asm volatile("" : "+v"(ls));
#pragma unroll
    for (int j = 0; j < kStreamChunk; ++j) {
      const uint32_t l = ls + j * kStep;
      stream_store(reinterpret_cast<V *>(psi + (base | seg[l >> L] | (l & segmask))),
                   tile_get16p<R>(ss0 ^ ((j & 1) ? sbit0 : 0u) ^ ((j & 2) ? sbit1 : 0u) ^ ((j & 4) ? sbit2 : 0u), spair));
    }
    __syncthreads();
    if (next >= n_tiles) break;  // uniform over the workgroup
#pragma unroll
    for (int j = 0; j < kStreamChunk; ++j) tile_put16<R>(s0 ^ slot_bytes<R>(j * kStep), v[j]);
    __syncthreads();
    base = nbase;
    const uint64_t cur = next;
    next = cur + gridDim.x;
    nbase = tile_base(next < n_tiles ? next : cur, P);
    if (next < n_tiles) prefetch(nbase);
  }
}

// Many circuits at once: workgroup b runs tile (b - job.first_block) of job block_job[b]; every
// job is one pass of one circuit whose state starts at psi + job.state_off.
template <typename R, int RB, bool TWO = true>
__global__ __launch_bounds__(kBlock) void k_batch_tile_pass(amp_t<R> *__restrict__ psi,
                                                            const BatchJob *__restrict__ jobs,
                                                            const uint32_t *__restrict__ block_job,
                                                            const TileStage *__restrict__ stages,
                                                            const RegOp *__restrict__ rops,
                                                            const TileOp *__restrict__ tops,
                                                            const amp_t<R> *__restrict__ mats) {
  extern __shared__ __align__(16) unsigned char smem[];
  const BatchJob *job = jobs + block_job[blockIdx.x];
  const TilePass &P = job->P;
  amp_t<R> *tile = reinterpret_cast<amp_t<R> *>(smem);
  uint64_t *seg = reinterpret_cast<uint64_t *>(smem + (sizeof(amp_t<R>) << P.t));
  fill_segments(P, seg);
  tile_body<R, RB, false, TWO>(psi + job->state_off, blockIdx.x - job->first_block, P, seg, tile, stages + job->stage_off,
                   rops + job->rop_off, tops + job->top_off, mats + job->mat_off);
}

// ------------------------------------------------------------------------------ batched circuits
template <typename R>
__global__ __launch_bounds__(kBlock) void k_batch_zero(amp_t<R> *__restrict__ psi, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    psi[i] = make_amp<R>(0.0, 0.0);
}
template <typename R>
__global__ void k_batch_ones(amp_t<R> *__restrict__ psi, const uint64_t *__restrict__ off, int n) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) psi[off[c]] = make_amp<R>(1.0, 0.0);
}

// ------------------------------------------------------------------------------ reductions
template <typename R>
__global__ __launch_bounds__(kBlock) void k_norm_masked(const amp_t<R> *__restrict__ psi,
                                                        uint64_t n_sub, BitInsert ins,
                                                        uint64_t val, double *partials) {
  __shared__ double lds[kBlock / 64];
  double acc = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < n_sub; g += stride) {
    amp_t<R> a = psi[insert_zeros(g, ins) | val];
    double x = re(a), y = im(a);
    acc = fma(x, x, fma(y, y, acc));
  }
  double t = block_sum(acc, lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ void k_finish_sum(const double *__restrict__ partials, int n_partials, int width,
                             double *out) {
  int j = threadIdx.x;
  if (j >= width) return;
  double t = 0.0;
  for (int b = 0; b < n_partials; ++b) t += partials[(size_t)b * width + j];
  out[j] = t;
}

// The measurement kernels run over a subspace: compact index c is the amplitude
// insert_zeros(c, sub) | sval, so the bits listed in `sub` are fixed to their values in sval (the
// caller knows every amplitude outside is zero).  With sub.count = 0 they cover the whole state.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_collapse(amp_t<R> *__restrict__ psi, uint64_t n_sub, BitInsert sub,
                                                     uint64_t sval, uint64_t mask, uint64_t val, double scale) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n_sub; c += stride) {
    const uint64_t i = insert_zeros(c, sub) | sval;
    amp_t<R> a = psi[i];
    if ((i & mask) != val) a = make_amp<R>(0.0, 0.0);
    else a = make_amp<R>(re(a) * scale, im(a) * scale);
    psi[i] = a;
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_chunk_norms(const amp_t<R> *__restrict__ psi, BitInsert sub,
                                                        uint64_t sval, int chunk_log2, double *sums) {
  __shared__ double lds[kBlock / 64];
  const uint64_t base = (uint64_t)blockIdx.x << chunk_log2;
  const uint64_t len = uint64_t(1) << chunk_log2;
  double acc = 0.0;
  for (uint64_t j = threadIdx.x; j < len; j += blockDim.x) {
    amp_t<R> a = psi[insert_zeros(base + j, sub) | sval];
    double x = re(a), y = im(a);
    acc = fma(x, x, fma(y, y, acc));
  }
  double t = block_sum(acc, lds);
  if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

// One workgroup: thread t owns the contiguous slice [t*per, (t+1)*per) of the chunk (compact
// indices of the subspace); the picked compact index goes to *pick.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_chunk_pick(const amp_t<R> *__restrict__ psi, BitInsert sub,
                                                       uint64_t sval, uint64_t start, uint64_t len, double base,
                                                       double rnd, unsigned long long *pick) {
  __shared__ double part[kBlock];
  const uint64_t per = (len + kBlock - 1) / kBlock;
  const uint64_t lo = min<uint64_t>(len, threadIdx.x * per), hi = min<uint64_t>(len, lo + per);
  double mine = 0.0;
  for (uint64_t j = lo; j < hi; ++j) {
    amp_t<R> a = psi[insert_zeros(start + j, sub) | sval];
    double x = re(a), y = im(a);
    mine += x * x + y * y;
  }
  part[threadIdx.x] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {  // exclusive prefix in slice order
    double run = base;
    for (int t = 0; t < kBlock; ++t) {
      double v = part[t];
      part[t] = run;
      run += v;
    }
  }
  __syncthreads();
  double run = part[threadIdx.x];
  for (uint64_t j = lo; j < hi; ++j) {
    amp_t<R> a = psi[insert_zeros(start + j, sub) | sval];
    double x = re(a), y = im(a);
    run += x * x + y * y;
    if (run >= rnd) {
      atomicMin(pick, (unsigned long long)(start + j));
      break;
    }
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_classical_flags(const amp_t<R> *__restrict__ psi,
                                                            uint64_t n, int p, double tol,
                                                            int *flags) {
  int lo = 0, hi = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    amp_t<R> a = psi[i];
    double x = re(a), y = im(a);
    int big = (x * x + y * y) > tol;
    if ((i >> p) & 1) hi |= big;
    else lo |= big;
  }
  lo = __syncthreads_or(lo);
  hi = __syncthreads_or(hi);
  if (threadIdx.x == 0) {
    if (lo) atomicOr(&flags[0], 1);
    if (hi) atomicOr(&flags[1], 1);
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_first_pair(const amp_t<R> *__restrict__ psi,
                                                       uint64_t n_pairs, int p, double tol,
                                                       unsigned long long *key) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  unsigned long long best = ~0ull;
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_pairs; q += stride) {
    const uint64_t i0 = insert_one_zero(q, p), i1 = i0 | (uint64_t(1) << p);
    amp_t<R> a = psi[i0], b = psi[i1];
    double n0 = re(a) * re(a) + im(a) * im(a), n1 = re(b) * re(b) + im(b) * im(b);
    if (n0 > tol) {
      best = min(best, (unsigned long long)(2 * q));
      break;  // q only grows in this thread
    }
    if (n1 > tol) {
      best = min(best, (unsigned long long)(2 * q + 1));
      break;
    }
  }
  if (best != ~0ull) atomicMin(key, best);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_drop_bit(const amp_t<R> *__restrict__ src,
                                                     amp_t<R> *__restrict__ dst, uint64_t n_out,
                                                     int p, int value) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t set = uint64_t(value ? 1 : 0) << p;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n_out; j += stride)
    dst[j] = src[insert_one_zero(j, p) | set];
}

// A qubit layout change, one exchange of two index bits at a time: the amplitudes whose bits a and
// b differ trade places (half the state read and written; a one-off when the layout is handed out).
template <typename R>
__global__ __launch_bounds__(kBlock) void k_swap_bits(amp_t<R> *__restrict__ psi, uint64_t n_pairs, int lo, int hi) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < n_pairs; c += stride) {
    const uint64_t i = insert_one_zero(insert_one_zero(c, lo), hi) | (uint64_t(1) << lo);
    const uint64_t j = i ^ (uint64_t(1) << lo) ^ (uint64_t(1) << hi);
    const amp_t<R> x = psi[i], y = psi[j];
    psi[i] = y;
    psi[j] = x;
  }
}

// The same layout change in one pass into another buffer, through LDS tiles so that both the
// reads and the writes run along contiguous segments (PermTile).  Every index is GF(2)-linear in
// its tile bits, so a thread adds one XOR per element to its own offsets.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_permute_tiles(const amp_t<R> *__restrict__ src, amp_t<R> *__restrict__ dst,
                                                          uint64_t n_tiles, PermTile P) {
  extern __shared__ __align__(16) unsigned char smem[];
  amp_t<R> *tile = reinterpret_cast<amp_t<R> *>(smem);
  constexpr int kTurns = (1 << kPermTileMaxBits) / kBlock;  // 16 turns of 256 elements at most
  const uint32_t size = 1u << P.m;
  auto spread = [&](uint32_t x, const uint8_t *pos) {
    uint64_t v = 0;
    for (uint32_t r = 0; r < P.m; ++r) v |= uint64_t((x >> r) & 1u) << pos[r];
    return v;
  };
  auto gather_store = [&](uint32_t y) {
    uint32_t x = 0;
    for (uint32_t r = 0; r < P.m; ++r) x |= ((y >> r) & 1u) << P.st_from[r];
    return x;
  };
  // Offsets are linear in the tile index: a thread's own part, plus the part of turn j (bits 8..11
  // of the tile index), the XOR of four uniform values.
  const uint64_t src_t = spread(threadIdx.x, P.src_pos), dst_t = spread(threadIdx.x, P.dst_pos);
  const uint32_t lds_t = gather_store(threadIdx.x);
  uint64_t src_j[4], dst_j[4];
  uint32_t lds_j[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    src_j[b] = spread(uint32_t(kBlock) << b, P.src_pos);
    dst_j[b] = spread(uint32_t(kBlock) << b, P.dst_pos);
    lds_j[b] = gather_store(uint32_t(kBlock) << b);
  }
  for (uint64_t tid = blockIdx.x; tid < n_tiles; tid += gridDim.x) {
    uint64_t sb = 0, db = 0;
    for (uint32_t i = 0; i < P.n_outer; ++i) {
      const uint64_t bit = (tid >> i) & 1u;
      sb |= bit << P.outer_src[i];
      db |= bit << P.outer_dst[i];
    }
    sb |= src_t;
    db |= dst_t;
#pragma unroll
    for (int j = 0; j < kTurns; ++j) {
      if ((uint32_t)j * kBlock + threadIdx.x < size) {
        const uint64_t o = ((j & 1) ? src_j[0] : 0) ^ ((j & 2) ? src_j[1] : 0) ^ ((j & 4) ? src_j[2] : 0) ^
                           ((j & 8) ? src_j[3] : 0);
        tile[j * kBlock + threadIdx.x] = src[sb ^ o];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kTurns; ++j) {
      if ((uint32_t)j * kBlock + threadIdx.x < size) {
        const uint64_t o = ((j & 1) ? dst_j[0] : 0) ^ ((j & 2) ? dst_j[1] : 0) ^ ((j & 4) ? dst_j[2] : 0) ^
                           ((j & 8) ? dst_j[3] : 0);
        const uint32_t x = lds_t ^ ((j & 1) ? lds_j[0] : 0u) ^ ((j & 2) ? lds_j[1] : 0u) ^ ((j & 4) ? lds_j[2] : 0u) ^
                           ((j & 8) ? lds_j[3] : 0u);
        dst[db ^ o] = tile[x];
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------ Pauli strings
// Pauli strings in expectation values are Hermitian (the caller drops the others, whose real part
// is zero): for x != 0 the amplitudes i and j = i ^ x contribute complex-conjugate terms to
// <psi|omega X^x Z^z|psi>, so each pair is visited once, from the i with bit h (the highest bit of
// x) clear, and counted twice.  That reads every amplitude once per string instead of twice.

// partials[2b + {0,1}] <- block sums of f (-1)^popcount(j & z) conj(psi_i) psi_j, j = i ^ x, over the
// i with bit h clear, f = 2 (x != 0); over every i with f = 1 (x = 0).  omega times the total has
// the real part of <psi|omega X^x Z^z|psi>.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_pauli_expect(const amp_t<R> *__restrict__ psi,
                                                         uint64_t n, uint64_t x, uint64_t z,
                                                         double *partials) {
  __shared__ double lds[kBlock / 64];
  Kahan sr, si;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  if (x == 0) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
      const amp_t<R> a = psi[i];
      const double v = (double)re(a) * re(a) + (double)im(a) * im(a);
      sr.add((__popcll(i & z) & 1) ? -v : v);
    }
  } else {
    const int hb = 63 - __clzll(x);
    const uint64_t lo = (uint64_t(1) << hb) - 1;
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < (n >> 1); g += stride) {
      const uint64_t i = ((g & ~lo) << 1) | (g & lo);
      const uint64_t j = i ^ x;
      const amp_t<R> a = psi[i], b = psi[j];
      const double f = (__popcll(j & z) & 1) ? -2.0 : 2.0;
      sr.add(f * ((double)re(a) * re(b) + (double)im(a) * im(b)));
      si.add(f * ((double)re(a) * im(b) - (double)im(a) * re(b)));
    }
  }
  double tr = block_sum(sr.value(), lds);
  double ti = block_sum(si.value(), lds);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = tr;
    partials[2 * blockIdx.x + 1] = ti;
  }
}

// All strings in one pass over a state that fits the caches: partials[b] <- block sums of
// Re(conj(psi_i) u_i), u_i = sum_t f_t c_t (-1)^popcount(j_t & z_t) psi_{j_t}, j_t = i ^ x_t, where a
// string with x_t != 0 counts only at the i with bit h_t clear, with f_t = 2 (f_t = 1 for x_t = 0).
// Partner loads are issued kInFlight strings at a time.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_pauli_energy(const amp_t<R> *__restrict__ psi, uint64_t n,
                                                         const PauliTerm *__restrict__ terms,
                                                         int n_terms, double *partials) {
  constexpr int kInFlight = 8;
  __shared__ double lds[kBlock / 64];
  Kahan acc;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const amp_t<R> a = psi[i];
    double ur1 = 0.0, ui1 = 0.0, ur2 = 0.0, ui2 = 0.0;  // strings with x = 0, paired strings
    auto term = [&](const PauliTerm &pt, const amp_t<R> &b) {
      double br = re(b), bi = im(b);
      if (__popcll((i ^ pt.x) & pt.z) & 1) {
        br = -br;
        bi = -bi;
      }
      const double vr = br * pt.c_re - bi * pt.c_im, vi = br * pt.c_im + bi * pt.c_re;
      if (pt.x) {
        ur2 += vr;
        ui2 += vi;
      } else {
        ur1 += vr;
        ui1 += vi;
      }
    };
    int t = 0;
    for (; t + kInFlight <= n_terms; t += kInFlight) {
      amp_t<R> b[kInFlight];
      bool on[kInFlight];
#pragma unroll
      for (int u = 0; u < kInFlight; ++u) {
        const uint64_t x = terms[t + u].x;
        on[u] = x == 0 || !((i >> (63 - __clzll(x))) & 1);
        if (on[u]) b[u] = psi[i ^ x];
      }
#pragma unroll
      for (int u = 0; u < kInFlight; ++u)
        if (on[u]) term(terms[t + u], b[u]);
    }
    for (; t < n_terms; ++t) {
      const PauliTerm &pt = terms[t];
      if (pt.x && ((i >> (63 - __clzll(pt.x))) & 1)) continue;
      term(pt, psi[i ^ pt.x]);
    }
    acc.add((double)re(a) * (ur1 + 2.0 * ur2) + (double)im(a) * (ui1 + 2.0 * ui2));  // Re(conj(a) * u)
  }
  double t = block_sum(acc.value(), lds);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

template <typename R, bool EVOLVE>
__global__ __launch_bounds__(kBlock) void k_pauli_sum(const amp_t<R> *__restrict__ src,
                                                      amp_t<R> *__restrict__ dst,
                                                      amp_t<R> *__restrict__ out, uint64_t n,
                                                      const PauliTerm *__restrict__ terms,
                                                      int n_terms, double f_re, double f_im,
                                                      uint64_t cmask, double *partials) {
  __shared__ double lds[kBlock / 64];
  double nrm = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    double ur = 0.0, ui = 0.0;
    for (int t = 0; t < n_terms; ++t) {
      const PauliTerm pt = terms[t];
      const uint64_t j = i ^ pt.x;
      amp_t<R> b = src[j];
      double br = re(b), bi = im(b);
      if (__popcll(j & pt.z) & 1) {
        br = -br;
        bi = -bi;
      }
      ur += br * pt.c_re - bi * pt.c_im;
      ui += br * pt.c_im + bi * pt.c_re;
    }
    if (EVOLVE) {
      double vr = ur * f_re - ui * f_im, vi = ur * f_im + ui * f_re;
      ur = vr;
      ui = vi;
    }
    dst[i] = make_amp<R>(ur, ui);
    if (EVOLVE && (i & cmask) == cmask) {
      amp_t<R> o = out[i];
      out[i] = make_amp<R>(re(o) + ur, im(o) + ui);
      nrm += ur * ur + ui * ui;
    }
  }
  if (EVOLVE) {
    double t = block_sum(nrm, lds);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
  }
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_scale_masked(amp_t<R> *__restrict__ psi, uint64_t n,
                                                         uint64_t cmask, double cr, double ci) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    if ((i & cmask) != cmask) continue;
    amp_t<R> a = psi[i];
    psi[i] = make_amp<R>(re(a) * cr - im(a) * ci, re(a) * ci + im(a) * cr);
  }
}

// ------------------------------------------------------------------------------ emulate_math
__device__ __forceinline__ int32_t wrap32(int64_t x) { return (int32_t)(uint32_t)(uint64_t)x; }
__device__ __forceinline__ int32_t check32(int64_t x, int &bad) {
  if (x < INT32_MIN || x > INT32_MAX) bad = 1;
  return (int32_t)x;
}
__device__ __forceinline__ int64_t floordiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}
__device__ __forceinline__ int64_t py_bitlen(int64_t x) {  // len(bin(x)[2:])
  if (x == 0) return 1;
  if (x > 0) return 64 - __clzll((unsigned long long)x);
  return 64 - __clzll((unsigned long long)(-x)) + 1;
}

// new values of the registers for opcode args.opcode (include/projectq_amd.h QS_MATH_*)
__device__ void math_eval(const MathArgs &A, int32_t *v, int &bad) {
  switch (A.opcode) {
    case 0:
      for (int r = 0; r < A.n_regs; ++r) v[r] = wrap32((int64_t)v[r] + (int32_t)A.a);
      break;
    case 1:
      for (int r = 0; r < A.n_regs; ++r)
        v[r] = wrap32(wrap32((int64_t)v[r] + (int32_t)A.a) % (int32_t)A.N);
      break;
    case 2:
      for (int r = 0; r < A.n_regs; ++r)
        v[r] = wrap32(wrap32((int64_t)v[r] * (int32_t)A.a) % (int32_t)A.N);
      break;
    case 3: {
      int64_t s = (int64_t)v[0] + v[1];
      if (py_bitlen(s) > (int64_t)A.width[0]) {
        int64_t mod = int64_t(1) << A.width[0];
        s = ((s % mod) + mod) % mod;
        if (A.n_regs == 3) v[2] ^= 1;
      }
      v[1] = check32(s, bad);
      break;
    }
    case 4:
      v[1] = check32((int64_t)v[1] - v[0], bad);
      break;
    case 5:
      if (A.n_regs == 3) v[2] ^= 1;
      v[1] = check32((int64_t)v[1] - v[0], bad);
      break;
    case 6:
      if (v[1] < v[0]) v[2] = (v[2] == 0) ? 1 : 0;
      break;
    case 7:
      v[2] = check32((int64_t)v[2] + (int64_t)v[0] * v[1], bad);
      break;
    case 8:
      v[2] = check32((int64_t)v[2] - (int64_t)v[0] * v[1], bad);
      break;
    case 9: {
      int64_t dividend = v[0], rem = v[1], divisor = v[2];
      if (divisor == 0 || divisor > dividend) {
        v[0] = check32(rem, bad);
        v[1] = check32(dividend, bad);
      } else {
        int64_t q = rem + floordiv(dividend, divisor);
        v[0] = check32(dividend - q * divisor, bad);
        v[1] = check32(q, bad);
      }
      break;
    }
    case 10: {
      int64_t rem = v[0], quo = v[1], divisor = v[2];
      if (divisor == 0) {
        v[0] = check32(quo, bad);
        v[1] = check32(rem, bad);
      } else {
        v[0] = check32(rem + quo * divisor, bad);
        v[1] = 0;
      }
      break;
    }
    default:
      bad = 1;
  }
}

// destination index of basis state i (simulator.hpp:246-265)
__device__ uint64_t math_dest(uint64_t i, const MathArgs &A, const int32_t *table, int &bad) {
  if ((i & A.cmask) != A.cmask) return i;
  int32_t v[kMaxRegs];
  uint64_t packed = 0;
  int shift = 0;
  for (int r = 0; r < A.n_regs; ++r) {
    uint64_t x = 0;
    for (int b = 0; b < A.width[r]; ++b) x |= ((i >> A.pos[r][b]) & 1ull) << b;
    v[r] = (int32_t)(uint32_t)x;
    packed |= x << shift;
    shift += A.width[r];
  }
  if (table) {
    for (int r = 0; r < A.n_regs; ++r) v[r] = table[packed * A.n_regs + r];
  } else {
    math_eval(A, v, bad);
  }
  uint64_t j = i;
  for (int r = 0; r < A.n_regs; ++r)
    for (int b = 0; b < A.width[r]; ++b) {
      uint64_t want = (uint64_t)(((int64_t)v[r] >> (b < 63 ? b : 63)) & 1);
      j = (j & ~(uint64_t(1) << A.pos[r][b])) | (want << A.pos[r][b]);
    }
  return j;
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_math_permute(const amp_t<R> *__restrict__ src,
                                                         amp_t<R> *__restrict__ dst, uint64_t n,
                                                         MathArgs A, const int32_t *table,
                                                         int *err) {
  int bad = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint64_t j = math_dest(i, A, table, bad);
    amp_t<R> a = src[i];
    dst[j] = make_amp<R>(0.0 + re(a), 0.0 + im(a));  // newvec starts at 0 (simulator.hpp:239-240)
  }
  if (bad) atomicOr(err, 1);
}

__global__ __launch_bounds__(kBlock) void k_math_injective(uint64_t n, MathArgs A,
                                                           const int32_t *table, uint32_t *bitmap,
                                                           int *collide, int *err) {
  int bad = 0, hit = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint64_t j = math_dest(i, A, table, bad);
    uint32_t bit = 1u << (j & 31);
    uint32_t old = atomicOr(&bitmap[j >> 5], bit);
    if (old & bit) hit = 1;
  }
  if (bad) atomicOr(err, 1);
  if (hit) atomicOr(collide, 1);
}

__global__ __launch_bounds__(kBlock) void k_math_keys(uint64_t n, MathArgs A, const int32_t *table,
                                                      uint32_t *keys, uint32_t *vals, int *err) {
  int bad = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    keys[i] = (uint32_t)math_dest(i, A, table, bad);
    vals[i] = (uint32_t)i;
  }
  if (bad) atomicOr(err, 1);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_math_segments(const amp_t<R> *__restrict__ src,
                                                          amp_t<R> *__restrict__ dst, uint64_t n,
                                                          const uint32_t *keys,
                                                          const uint32_t *vals) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += stride) {
    if (j > 0 && keys[j] == keys[j - 1]) continue;
    double sr = 0.0, si = 0.0;
    uint64_t k = j;
    do {
      amp_t<R> a = src[vals[k]];
      sr += re(a);
      si += im(a);
      ++k;
    } while (k < n && keys[k] == keys[j]);
    dst[keys[j]] = make_amp<R>(sr, si);
  }
}

template <typename Ra, typename Rb>
__global__ __launch_bounds__(kBlock) void k_overlap(const amp_t<Ra> *__restrict__ a,
                                                    const amp_t<Rb> *__restrict__ b, uint64_t n,
                                                    double *partials) {
  __shared__ double lds[kBlock / 64];
  Kahan sr, si;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    amp_t<Ra> x = a[i];
    amp_t<Rb> y = b[i];
    sr.add(re(x) * re(y) + im(x) * im(y));
    si.add(re(x) * im(y) - im(x) * re(y));
  }
  double tr = block_sum(sr.value(), lds);
  double ti = block_sum(si.value(), lds);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = tr;
    partials[2 * blockIdx.x + 1] = ti;
  }
}

}  // namespace

// ------------------------------------------------------------------------------ launchers
void overlap(const void *a, bool a_c64, const void *b, bool b_c64, uint64_t n, double *partials,
             hipStream_t s) {
  dim3 g(kReduceGrid), t(kBlock);
  if (!a_c64 && !b_c64)
    hipLaunchKernelGGL((k_overlap<double, double>), g, t, 0, s, (const double2 *)a, (const double2 *)b, n, partials);
  else if (!a_c64 && b_c64)
    hipLaunchKernelGGL((k_overlap<double, float>), g, t, 0, s, (const double2 *)a, (const float2 *)b, n, partials);
  else if (a_c64 && !b_c64)
    hipLaunchKernelGGL((k_overlap<float, double>), g, t, 0, s, (const float2 *)a, (const double2 *)b, n, partials);
  else
    hipLaunchKernelGGL((k_overlap<float, float>), g, t, 0, s, (const float2 *)a, (const float2 *)b, n, partials);
  QS_HIP(hipGetLastError());
}
template <typename R>
void apply_block(amp_t<R> *psi, uint64_t n_groups, const BitInsert &ins, uint64_t cmask,
                 const uint32_t *tpos, int k, const amp_t<R> *d_mat, hipStream_t s) {
  switch (k) {
    case 1: launch_apply<R, 1>(psi, n_groups, ins, cmask, tpos, d_mat, s); break;
    case 2: launch_apply<R, 2>(psi, n_groups, ins, cmask, tpos, d_mat, s); break;
    case 3: launch_apply<R, 3>(psi, n_groups, ins, cmask, tpos, d_mat, s); break;
    case 4: launch_apply<R, 4>(psi, n_groups, ins, cmask, tpos, d_mat, s); break;
    case 5: launch_apply<R, 5>(psi, n_groups, ins, cmask, tpos, d_mat, s); break;
    default: throw ValueError("Gates with more than 5 qubits are not supported!");
  }
}

template <typename R, int RB>
void launch_tile_pass(amp_t<R> *psi, uint64_t n_tiles, const TilePass &P, const TileStage *d_stages,
                      const RegOp *d_rops, const TileOp *d_tops, const amp_t<R> *d_mats, bool two_qubit_ops,
                      hipStream_t s) {
  prepare_tile_kernels();
  const size_t lds = (sizeof(amp_t<R>) << P.t) + (sizeof(uint64_t) << P.h);
  // workgroups per launch: each loops over n_tiles / grid tiles.  About 16 tiles per workgroup
  // measured best at 2^20 tiles (18.5 ms per pass against 19.1 ms at 128 per workgroup and 19.8
  // ms at one), never fewer than 8192 workgroups; PROJECTQ_AMD_TILE_GRID fixes the count.
  static const uint64_t cap = [] {
    const char *e = std::getenv("PROJECTQ_AMD_TILE_GRID");
    return e ? std::max<uint64_t>(1, std::strtoull(e, nullptr, 10)) : uint64_t(0);
  }();
  // the pipelined kernel is the default where it applies; PROJECTQ_AMD_TILE_PREFETCH=0 turns it off
  static const bool prefetch = [] {
    const char *e = std::getenv("PROJECTQ_AMD_TILE_PREFETCH");
    return !e || std::atoi(e) != 0;
  }();
  const uint64_t grid = std::min<uint64_t>(n_tiles, cap ? cap : std::max<uint64_t>(8192, n_tiles / 16));
  const uint64_t chunk = uint64_t(kStreamChunk) * kBlock * Vec16<R>::n;  // amplitudes streamed per round
  const uint64_t tsize = uint64_t(1) << P.t;
  bool done = false;
  if constexpr (RB == 4 && sizeof(R) == 4) {  // the only instantiations of the pipelined kernel
    if (prefetch && tsize == chunk) {
      if (two_qubit_ops)
        hipLaunchKernelGGL((k_tile_pass_pf<R, RB, kBlock, true>), dim3((unsigned)grid), dim3(kBlock), lds, s, psi,
                           n_tiles, P, d_stages, d_rops, d_tops, d_mats);
      else
        hipLaunchKernelGGL((k_tile_pass_pf<R, RB, kBlock, false>), dim3((unsigned)grid), dim3(kBlock), lds, s, psi,
                           n_tiles, P, d_stages, d_rops, d_tops, d_mats);
      done = true;
    } else if (prefetch && tsize == 2 * chunk) {
      if (two_qubit_ops)
        hipLaunchKernelGGL((k_tile_pass_pf<R, RB, 2 * kBlock, true>), dim3((unsigned)grid), dim3(2 * kBlock), lds, s,
                           psi, n_tiles, P, d_stages, d_rops, d_tops, d_mats);
      else
        hipLaunchKernelGGL((k_tile_pass_pf<R, RB, 2 * kBlock, false>), dim3((unsigned)grid), dim3(2 * kBlock), lds,
                           s, psi, n_tiles, P, d_stages, d_rops, d_tops, d_mats);
      done = true;
    }
  }
  if constexpr (RB == 4) {  // passes without 2-qubit register ops: the kernel without those forms
    if (!done && tsize % chunk == 0 && !two_qubit_ops) {
      hipLaunchKernelGGL((k_tile_pass<R, RB, true, false>), dim3((unsigned)grid), dim3(kBlock), lds, s, psi,
                         n_tiles, P, d_stages, d_rops, d_tops, d_mats);
      done = true;
    }
  }
  if (done) {
  } else if (tsize % chunk == 0)
    hipLaunchKernelGGL((k_tile_pass<R, RB, true>), dim3((unsigned)grid), dim3(kBlock), lds, s, psi, n_tiles, P,
                       d_stages, d_rops, d_tops, d_mats);
  else
    hipLaunchKernelGGL((k_tile_pass<R, RB, false>), dim3((unsigned)grid), dim3(kBlock), lds, s, psi, n_tiles, P,
                       d_stages, d_rops, d_tops, d_mats);
  QS_HIP(hipGetLastError());
}

template <typename R>
void tile_pass(amp_t<R> *psi, uint64_t n_tiles, const TilePass &P, int rb, const TileStage *d_stages,
               const RegOp *d_rops, const TileOp *d_tops, const amp_t<R> *d_mats, bool two_qubit_ops, hipStream_t s) {
  if (P.h > (uint32_t)kTileMaxHigh || P.L < 1 || P.t > (uint32_t)kTileMaxBits)
    throw ValueError("tile pass: bad tile shape");
  if (rb == 5) launch_tile_pass<R, 5>(psi, n_tiles, P, d_stages, d_rops, d_tops, d_mats, two_qubit_ops, s);
  else if (rb == 4) launch_tile_pass<R, 4>(psi, n_tiles, P, d_stages, d_rops, d_tops, d_mats, two_qubit_ops, s);
  else throw ValueError("tile pass: register bits must be 4 or 5");
}

// largest dynamic LDS of a tile kernel: the biggest tile plus the biggest segment table
template <typename R>
constexpr int kTileLdsMax = int(sizeof(amp_t<R>) << kTileMaxBits) + int(sizeof(uint64_t) << kTileMaxHigh);

void prepare_tile_kernels() {
  static bool done = false;
  if (done) return;
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass<float, 4, true>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass<float, 5, true>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass<double, 4, true>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<double>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass<double, 5, true>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<double>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass<float, 4, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass<double, 4, true, false>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<double>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass<float, 4, false>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass<float, 5, false>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass<double, 4, false>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<double>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass<double, 5, false>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<double>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass_pf<float, 4, kBlock, true>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass_pf<float, 4, 2 * kBlock, true>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass_pf<float, 4, kBlock, false>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_tile_pass_pf<float, 4, 2 * kBlock, false>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_batch_tile_pass<float, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_batch_tile_pass<float, 5>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_batch_tile_pass<double, 4>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<double>));
  QS_HIP(hipFuncSetAttribute((const void *)k_batch_tile_pass<double, 5>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<double>));
  QS_HIP(hipFuncSetAttribute((const void *)k_batch_tile_pass<float, 4, false>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<float>));
  QS_HIP(hipFuncSetAttribute((const void *)k_batch_tile_pass<double, 4, false>, hipFuncAttributeMaxDynamicSharedMemorySize, kTileLdsMax<double>));
  done = true;
}

template <typename R>
void batch_init(amp_t<R> *psi, uint64_t total, const uint64_t *state_off, int n_states, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_zero<R>, dim3(grid_for(total)), dim3(kBlock), 0, s, psi, total);
  QS_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_batch_ones<R>, dim3((n_states + 255) / 256), dim3(256), 0, s, psi, state_off, n_states);
  QS_HIP(hipGetLastError());
}

template <typename R, int RB>
void launch_batch_tile_pass(amp_t<R> *psi, const BatchJob *jobs, const uint32_t *block_job, uint32_t n_blocks,
                            uint32_t max_t, const TileStage *stages, const RegOp *rops, const TileOp *tops,
                            const amp_t<R> *mats, bool two_qubit_ops, hipStream_t s) {
  prepare_tile_kernels();
  const size_t lds = (sizeof(amp_t<R>) << max_t) + (sizeof(uint64_t) << kTileMaxHigh);
  bool done = false;
  if constexpr (RB == 4) {
    if (!two_qubit_ops) {
      hipLaunchKernelGGL((k_batch_tile_pass<R, RB, false>), dim3(n_blocks), dim3(kBlock), lds, s, psi, jobs,
                         block_job, stages, rops, tops, mats);
      done = true;
    }
  }
  if (!done)
    hipLaunchKernelGGL((k_batch_tile_pass<R, RB>), dim3(n_blocks), dim3(kBlock), lds, s, psi, jobs, block_job, stages,
                       rops, tops, mats);
  QS_HIP(hipGetLastError());
}

template <typename R>
void batch_tile_pass(amp_t<R> *psi, const BatchJob *jobs, const uint32_t *block_job, uint32_t n_blocks,
                     uint32_t max_t, int rb, const TileStage *stages, const RegOp *rops, const TileOp *tops,
                     const amp_t<R> *mats, bool two_qubit_ops, hipStream_t s) {
  if (max_t > (uint32_t)kTileMaxBits || n_blocks == 0) throw ValueError("batch pass: bad launch shape");
  if (rb == 5)
    launch_batch_tile_pass<R, 5>(psi, jobs, block_job, n_blocks, max_t, stages, rops, tops, mats, two_qubit_ops, s);
  else if (rb == 4)
    launch_batch_tile_pass<R, 4>(psi, jobs, block_job, n_blocks, max_t, stages, rops, tops, mats, two_qubit_ops, s);
  else throw ValueError("batch pass: register bits must be 4 or 5");
}

template <typename R>
void norm_masked(const amp_t<R> *psi, uint64_t n_sub, const BitInsert &ins, uint64_t val,
                 double *partials, hipStream_t s) {
  hipLaunchKernelGGL(k_norm_masked<R>, dim3(kReduceGrid), dim3(kBlock), 0, s, psi, n_sub, ins, val,
                     partials);
  QS_HIP(hipGetLastError());
}

void finish_sum(const double *partials, int n_partials, int width, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_finish_sum, dim3(1), dim3(64), 0, s, partials, n_partials, width, out);
  QS_HIP(hipGetLastError());
}

template <typename R>
void collapse(amp_t<R> *psi, uint64_t n_sub, const BitInsert &sub, uint64_t sval, uint64_t mask, uint64_t val,
              double scale, hipStream_t s) {
  hipLaunchKernelGGL(k_collapse<R>, dim3(grid_for(n_sub)), dim3(kBlock), 0, s, psi, n_sub, sub, sval, mask, val,
                     scale);
  QS_HIP(hipGetLastError());
}

template <typename R>
void chunk_norms(const amp_t<R> *psi, uint64_t n_sub, const BitInsert &sub, uint64_t sval, int chunk_log2,
                 double *sums, hipStream_t s) {
  uint64_t chunks = n_sub >> chunk_log2;
  hipLaunchKernelGGL(k_chunk_norms<R>, dim3((unsigned)chunks), dim3(kBlock), 0, s, psi, sub, sval, chunk_log2,
                     sums);
  QS_HIP(hipGetLastError());
}

template <typename R>
void chunk_pick(const amp_t<R> *psi, const BitInsert &sub, uint64_t sval, uint64_t start, uint64_t len, double base,
                double rnd, unsigned long long *pick, hipStream_t s) {
  hipLaunchKernelGGL(k_chunk_pick<R>, dim3(1), dim3(kBlock), 0, s, psi, sub, sval, start, len, base, rnd, pick);
  QS_HIP(hipGetLastError());
}

template <typename R>
void classical_flags(const amp_t<R> *psi, uint64_t n, int p, double tol, int *flags, hipStream_t s) {
  hipLaunchKernelGGL(k_classical_flags<R>, dim3(grid_for(n)), dim3(kBlock), 0, s, psi, n, p, tol,
                     flags);
  QS_HIP(hipGetLastError());
}

template <typename R>
void first_pair(const amp_t<R> *psi, uint64_t n_pairs, int p, double tol, unsigned long long *key,
                hipStream_t s) {
  hipLaunchKernelGGL(k_first_pair<R>, dim3(grid_for(n_pairs)), dim3(kBlock), 0, s, psi, n_pairs, p,
                     tol, key);
  QS_HIP(hipGetLastError());
}

template <typename R>
void drop_bit(const amp_t<R> *src, amp_t<R> *dst, uint64_t n_out, int p, int value, hipStream_t s) {
  hipLaunchKernelGGL(k_drop_bit<R>, dim3(grid_for(n_out)), dim3(kBlock), 0, s, src, dst, n_out, p,
                     value);
  QS_HIP(hipGetLastError());
}

template <typename R>
void swap_bits(amp_t<R> *psi, uint64_t n_amps, int a, int b, hipStream_t s) {
  if (a == b || n_amps < 4) return;
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  hipLaunchKernelGGL(k_swap_bits<R>, dim3(grid_for(n_amps / 4)), dim3(kBlock), 0, s, psi, n_amps / 4, lo, hi);
  QS_HIP(hipGetLastError());
}

template <typename R>
void permute_tiles(const amp_t<R> *src, amp_t<R> *dst, uint64_t n_amps, const PermTile &P, hipStream_t s) {
  if (P.m > (uint32_t)kPermTileMaxBits || (uint64_t(1) << P.m) > n_amps || P.m + P.n_outer > 64)
    throw ValueError("permute_tiles: bad tile");
  const uint64_t n_tiles = n_amps >> P.m;
  const uint64_t grid = std::min<uint64_t>(n_tiles, 16384);
  const size_t lds = sizeof(amp_t<R>) << P.m;
  hipLaunchKernelGGL(k_permute_tiles<R>, dim3((unsigned)grid), dim3(kBlock), lds, s, src, dst, n_tiles, P);
  QS_HIP(hipGetLastError());
}

template <typename R>
void pauli_expect(const amp_t<R> *psi, uint64_t n, uint64_t x, uint64_t z, double *partials,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_pauli_expect<R>, dim3(kReduceGrid), dim3(kBlock), 0, s, psi, n, x, z,
                     partials);
  QS_HIP(hipGetLastError());
}

template <typename R>
void pauli_energy(const amp_t<R> *psi, uint64_t n, const PauliTerm *terms, int n_terms, double *partials,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_pauli_energy<R>, dim3(kReduceGrid), dim3(kBlock), 0, s, psi, n, terms, n_terms,
                     partials);
  QS_HIP(hipGetLastError());
}

template <typename R>
void pauli_sum(const amp_t<R> *src, amp_t<R> *dst, amp_t<R> *out, uint64_t n,
               const PauliTerm *terms, int n_terms, double f_re, double f_im, bool evolve,
               uint64_t cmask, double *partials, hipStream_t s) {
  if (evolve)
    hipLaunchKernelGGL((k_pauli_sum<R, true>), dim3(kReduceGrid), dim3(kBlock), 0, s, src, dst, out,
                       n, terms, n_terms, f_re, f_im, cmask, partials);
  else
    hipLaunchKernelGGL((k_pauli_sum<R, false>), dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, out,
                       n, terms, n_terms, f_re, f_im, cmask, partials);
  QS_HIP(hipGetLastError());
}

template <typename R>
void scale_masked(amp_t<R> *psi, uint64_t n, uint64_t cmask, double cr, double ci, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_masked<R>, dim3(grid_for(n)), dim3(kBlock), 0, s, psi, n, cmask, cr, ci);
  QS_HIP(hipGetLastError());
}

template <typename R>
void math_permute(const amp_t<R> *src, amp_t<R> *dst, uint64_t n, const MathArgs &args,
                  const int32_t *table, int *err, hipStream_t s) {
  hipLaunchKernelGGL(k_math_permute<R>, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, n, args,
                     table, err);
  QS_HIP(hipGetLastError());
}

void math_injective(uint64_t n, const MathArgs &args, const int32_t *table, uint32_t *bitmap,
                    int *collide, int *err, hipStream_t s) {
  hipLaunchKernelGGL(k_math_injective, dim3(grid_for(n)), dim3(kBlock), 0, s, n, args, table, bitmap,
                     collide, err);
  QS_HIP(hipGetLastError());
}

template <typename R>
void math_accumulate(const amp_t<R> *src, amp_t<R> *dst, uint64_t n, const MathArgs &args,
                     const int32_t *table, int *err, hipStream_t s) {
  if (n > (uint64_t(1) << 30))
    throw ValueError("emulate_math: a non-injective function on more than 30 qubits is not supported");
  uint32_t *buf = nullptr;
  QS_HIP(hipMalloc(&buf, 4 * n * sizeof(uint32_t)));
  uint32_t *k_in = buf, *v_in = buf + n, *k_out = buf + 2 * n, *v_out = buf + 3 * n;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  int bits = 1;
  while ((uint64_t(1) << bits) < n) ++bits;
  try {
    hipLaunchKernelGGL(k_math_keys, dim3(grid_for(n)), dim3(kBlock), 0, s, n, args, table, k_in,
                       v_in, err);
    QS_HIP(hipGetLastError());
    QS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k_in, k_out, v_in, v_out, (int)n, 0,
                                              bits, s));
    QS_HIP(hipMalloc(&tmp, tmp_bytes));
    QS_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, k_in, k_out, v_in, v_out, (int)n, 0,
                                              bits, s));
    hipLaunchKernelGGL(k_math_segments<R>, dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, n, k_out,
                       v_out);
    QS_HIP(hipGetLastError());
    QS_HIP(hipStreamSynchronize(s));
  } catch (...) {
    (void)hipStreamSynchronize(s);
    (void)hipFree(tmp);
    (void)hipFree(buf);
    throw;
  }
  QS_HIP(hipFree(tmp));
  QS_HIP(hipFree(buf));
}

// ------------------------------------------------------------------------------ instantiations
template <typename S, typename D>
__global__ __launch_bounds__(kBlock) void k_convert_amps(const amp_t<S> *__restrict__ src, amp_t<D> *__restrict__ dst,
                                                         uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const amp_t<S> a = src[i];
    amp_t<D> b;
    b.x = (D)a.x;
    b.y = (D)a.y;
    dst[i] = b;
  }
}

template <typename S, typename D>
void convert_amps(const amp_t<S> *src, amp_t<D> *dst, uint64_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL((k_convert_amps<S, D>), dim3(grid_for(n)), dim3(kBlock), 0, s, src, dst, n);
  QS_HIP(hipGetLastError());
}
template void convert_amps<double, float>(const amp_t<double> *, amp_t<float> *, uint64_t, hipStream_t);
template void convert_amps<float, double>(const amp_t<float> *, amp_t<double> *, uint64_t, hipStream_t);

#define QS_INSTANTIATE(R)                                                                          \
  template void batch_init<R>(amp_t<R> *, uint64_t, const uint64_t *, int, hipStream_t);          \
  template void batch_tile_pass<R>(amp_t<R> *, const BatchJob *, const uint32_t *, uint32_t,       \
                                   uint32_t, int, const TileStage *, const RegOp *, const TileOp *,\
                                   const amp_t<R> *, bool, hipStream_t);                                \
  template void tile_pass<R>(amp_t<R> *, uint64_t, const TilePass &, int, const TileStage *,      \
                             const RegOp *, const TileOp *, const amp_t<R> *, bool, hipStream_t); \
  template void apply_block<R>(amp_t<R> *, uint64_t, const BitInsert &, uint64_t,                 \
                               const uint32_t *, int, const amp_t<R> *, hipStream_t);             \
  template void norm_masked<R>(const amp_t<R> *, uint64_t, const BitInsert &, uint64_t, double *, \
                               hipStream_t);                                                      \
  template void collapse<R>(amp_t<R> *, uint64_t, const BitInsert &, uint64_t, uint64_t, uint64_t, double,    \
                            hipStream_t);                                                         \
  template void chunk_norms<R>(const amp_t<R> *, uint64_t, const BitInsert &, uint64_t, int, double *,        \
                               hipStream_t);                                                      \
  template void chunk_pick<R>(const amp_t<R> *, const BitInsert &, uint64_t, uint64_t, uint64_t, double,      \
                              double, unsigned long long *, hipStream_t);                         \
  template void classical_flags<R>(const amp_t<R> *, uint64_t, int, double, int *, hipStream_t);  \
  template void first_pair<R>(const amp_t<R> *, uint64_t, int, double, unsigned long long *,      \
                              hipStream_t);                                                       \
  template void drop_bit<R>(const amp_t<R> *, amp_t<R> *, uint64_t, int, int, hipStream_t);       \
  template void swap_bits<R>(amp_t<R> *, uint64_t, int, int, hipStream_t);                          \
  template void permute_tiles<R>(const amp_t<R> *, amp_t<R> *, uint64_t, const PermTile &, hipStream_t); \
  template void pauli_expect<R>(const amp_t<R> *, uint64_t, uint64_t, uint64_t, double *,         \
                                hipStream_t);                                                     \
  template void pauli_energy<R>(const amp_t<R> *, uint64_t, const PauliTerm *, int, double *,      \
                                hipStream_t);                                                     \
  template void pauli_sum<R>(const amp_t<R> *, amp_t<R> *, amp_t<R> *, uint64_t,                  \
                             const PauliTerm *, int, double, double, bool, uint64_t, double *,    \
                             hipStream_t);                                                        \
  template void scale_masked<R>(amp_t<R> *, uint64_t, uint64_t, double, double, hipStream_t);     \
  template void math_permute<R>(const amp_t<R> *, amp_t<R> *, uint64_t, const MathArgs &,         \
                                const int32_t *, int *, hipStream_t);                             \
  template void math_accumulate<R>(const amp_t<R> *, amp_t<R> *, uint64_t, const MathArgs &,      \
                                   const int32_t *, int *, hipStream_t);

QS_INSTANTIATE(double)
QS_INSTANTIATE(float)

}  // namespace dev
}  // namespace qs
