// CPU emulation of the MFMA dense-block pass (projectq_amd/csrc/mfma.hip) for the tests: the
// matrix table of mfma_layout.hpp, the kernel's per-lane loads and stores ((re, im) pairs of one
// half of one group per lane), and the documented lane semantics of v_mfma_f32_32x32x2f32, in FP64.
#include <complex>
#include <cstdint>
#include <vector>

#include "mfma_layout.hpp"

using cplx = std::complex<double>;

extern "C" {

// groups: n_groups x D complex amplitudes (interleaved), updated in place; m: D x D row-major.
void mfma_emulate(int k, const double *m_il, double *groups, int n_groups) {
  const int D = 1 << k, H = D / 2, NB = qs::mfma::blocks(k), S = qs::mfma::steps(k);
  const int QQ = k == 3 ? 2 : 4;
  std::vector<cplx> m(D * D);
  for (int i = 0; i < D * D; ++i) m[i] = cplx(m_il[2 * i], m_il[2 * i + 1]);
  const std::vector<float> table = qs::mfma::operand_table(k, m.data());
  auto comp = [&](int g, int a, int cc) -> double & { return groups[(size_t)(g * D + a) * 2 + cc]; };
  for (int c0 = 0; c0 < n_groups; c0 += 32) {
    // lane loads: x[L][m] = (re, im) of amplitude H h + m of group n
    std::vector<std::vector<double>> xr(64, std::vector<double>(H)), xi(64, std::vector<double>(H));
    for (int L = 0; L < 64; ++L) {
      const int n = L & 31, h = L >> 5, g = c0 + n;
      for (int mm = 0; mm < H; ++mm) {
        xr[L][mm] = g < n_groups ? comp(g, H * h + mm, 0) : 0.0;
        xi[L][mm] = g < n_groups ? comp(g, H * h + mm, 1) : 0.0;
      }
    }
    std::vector<std::vector<double>> out(64, std::vector<double>(NB * 16, 0.0));
    for (int b = 0; b < NB; ++b) {
      double acc[64][16] = {};
      for (int i = 0; i < S; ++i) {
        // B operand of step i: lane L gives component i % 2 of its amplitude i / 2
        auto bval = [&](int L) { return (i & 1) ? xi[L][i >> 1] : xr[L][i >> 1]; };
        // 32x32x2: A[r][kk] from lane r + 32 kk, B[kk][j] from lane j + 32 kk, D[r][j] in lane
        // j + 32 ((r / 4) % 2), register 4 (r / 8) + r % 4
        for (int r = 0; r < 32; ++r)
          for (int j = 0; j < 32; ++j) {
            double sum = 0.0;
            for (int kk = 0; kk < 2; ++kk) sum += (double)table[(b * S + i) * 64 + r + 32 * kk] * bval(j + 32 * kk);
            acc[j + 32 * ((r / 4) % 2)][4 * (r / 8) + r % 4] += sum;
          }
      }
      for (int L = 0; L < 64; ++L)
        for (int j = 0; j < 16; ++j) out[L][16 * b + j] = acc[L][j];
    }
    for (int L = 0; L < 64; ++L) {
      const int n = L & 31, h = L >> 5, g = c0 + n;
      if (g >= n_groups) continue;
      for (int b = 0; b < NB; ++b)
        for (int qq = 0; qq < QQ; ++qq)
          for (int s2 = 0; s2 < 2; ++s2) {
            const int a = H * h + 8 * b + 2 * qq + s2;
            comp(g, a, 0) = out[L][16 * b + 4 * qq + 2 * s2];
            comp(g, a, 1) = out[L][16 * b + 4 * qq + 2 * s2 + 1];
          }
    }
  }
}
}
