// Operand layout of the MFMA dense-block pass (mfma.hip), shared by the kernel, the host code that
// builds the matrix table, and the CPU emulation in the tests.
//
// A k-qubit block M (2^k x 2^k complex, row = output) applied to a group of D = 2^k amplitudes is
// the real product  y = A x  with x, y the 2D real components (re, im) of the group and
//   A[(ao, re)][(ai, re)] =  Re M[ao][ai]     A[(ao, re)][(ai, im)] = -Im M[ao][ai]
//   A[(ao, im)][(ai, re)] =  Im M[ao][ai]     A[(ao, im)][(ai, im)] =  Re M[ao][ai]
// Batched over groups (the columns of X and Y) it is a GEMM with M = K = 2D on
// v_mfma_f32_32x32x2f32: the matrix is the A operand, each lane's amplitudes the B operand, the
// results come back in the C registers.
//
// Lane L serves group n = L & 31 and half h = L >> 5 of it: amplitudes H h + m, m < H = D / 2
// (the half whose top matrix bit is h), which it loads and stores as (re, im) pairs.  The row
// orders below make that possible without moving data between lanes:
//   B, step i (i < D):                 row 2i + h = amplitude H h + i / 2, component i % 2
//   A, block b (< NB), step i:          A[32b + (L & 31)][2i + (L >> 5)]
//   C, block b, register j = 4qq + s:   row 32b + 8qq + 4h + s
//                                       = amplitude H h + 8b + 2qq + s / 2, component s % 2
// NB = D / 16 output blocks of 32 rows (k = 4, 5); for k = 3 one block whose rows with qq >= 2
// are zero (2D = 16 real rows).
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

namespace qs {
namespace mfma {

constexpr int blocks(int k) { return k >= 4 ? (1 << k) / 16 : 1; }
constexpr int steps(int k) { return 1 << k; }
// table floats: one per lane, output block and step
constexpr int table_size(int k) { return 64 * blocks(k) * steps(k); }

// (amplitude, component) of the A matrix's output row R; false for the zero rows of k = 3
inline bool out_row(int k, int R, int *amp, int *comp) {
  const int H = (1 << k) / 2;
  const int b = R >> 5, qq = (R >> 3) & 3, h = (R >> 2) & 1, s = R & 3;
  if (k == 3 && qq >= 2) return false;
  *amp = H * h + 8 * b + 2 * qq + (s >> 1);
  *comp = s & 1;
  return true;
}
// (amplitude, component) of the A matrix's input column c = 2i + h
inline void in_row(int k, int c, int *amp, int *comp) {
  const int H = (1 << k) / 2, i = c >> 1, h = c & 1;
  *amp = H * h + (i >> 1);
  *comp = i & 1;
}

// A[(ao, co)][(ai, ci)] of the real form of M (row-major D x D complex)
inline double real_entry(const std::complex<double> *m, int D, int ao, int co, int ai, int ci) {
  const std::complex<double> v = m[ao * D + ai];
  if (co == ci) return v.real();
  return co == 0 ? -v.imag() : v.imag();
}

// The A-operand table, entry-major: t[(b * steps + i) * 64 + L] is lane L's A operand for output
// block b and step i.
inline std::vector<float> operand_table(int k, const std::complex<double> *m) {
  const int D = 1 << k, NB = blocks(k), S = steps(k);
  std::vector<float> t(table_size(k), 0.0f);
  for (int b = 0; b < NB; ++b)
    for (int i = 0; i < S; ++i)
      for (int L = 0; L < 64; ++L) {
        int ao, co, ai, ci;
        if (!out_row(k, 32 * b + (L & 31), &ao, &co)) continue;
        in_row(k, 2 * i + (L >> 5), &ai, &ci);
        t[(b * S + i) * 64 + L] = (float)real_entry(m, D, ao, co, ai, ci);
      }
  return t;
}

}  // namespace mfma
}  // namespace qs
