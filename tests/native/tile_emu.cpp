// Test-only CPU model of the tile-pass kernel (k_tile_pass in projectq_amd/csrc/kernels.hip).
//
// Runs the product planner (planner.cpp: merge, schedule, build_tile_pass) on a gate list and
// replays every tile pass on the CPU with the kernel's own index arithmetic, checking each
// global and LDS index against its bounds.  Lets the CPU test-suite check that the planner's
// passes are in bounds and that their result equals the unfused gate sequence.
#include <complex>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "planner.hpp"

using cplx = std::complex<double>;

namespace {

uint64_t insert_zeros(uint64_t g, const qs::BitInsert &ins) {
  for (uint32_t i = 0; i < ins.count; ++i) {
    uint32_t p = ins.pos[i];
    g = ((g >> p) << (p + 1)) | (g & ((uint64_t(1) << p) - 1));
  }
  return g;
}

struct Fail {
  std::string what;
};

void check(bool ok, const char *what) {
  if (!ok) throw Fail{what};
}

void replay(cplx *psi, unsigned n, const qs::dev::TilePass &P, const std::vector<qs::dev::TileOp> &ops,
            const std::vector<cplx> &mats) {
  const uint32_t t = P.t, L = P.L, h = P.h;
  const uint64_t N = uint64_t(1) << n;
  const uint32_t tsize = 1u << t, segmask = (1u << L) - 1;
  check(t <= n && L + h == t && h <= (uint32_t)qs::dev::kTileMaxHigh, "tile shape");
  std::vector<uint64_t> seg(1u << h);
  for (uint32_t s = 0; s < (1u << h); ++s) {
    uint64_t o = 0;
    for (uint32_t b = 0; b < h; ++b)
      if ((s >> b) & 1) o |= uint64_t(1) << P.hpos[b];
    seg[s] = o;
  }
  const uint64_t n_tiles = N >> t;
  std::vector<cplx> tile(tsize);
  for (uint64_t tix = 0; tix < n_tiles; ++tix) {
    const uint64_t base = insert_zeros(tix, P.tiles);
    check(base < N, "tile base");
    for (uint32_t l = 0; l < tsize; ++l) {
      check((l >> L) < seg.size(), "segment index");
      uint64_t gi = base | seg[l >> L] | (l & segmask);
      check(gi < N, "global index");
      tile[l] = psi[gi];
    }
    for (const auto &op : ops) {
      if ((base & op.gcmask) != op.gcmask) continue;
      const uint32_t D = 1u << op.k;
      check(op.k >= 1 && op.k <= 5 && op.nfixed <= t, "op shape");
      check(op.mat + D * D <= mats.size(), "matrix offset");
      const uint32_t ngroups = 1u << (t - op.nfixed);
      std::vector<cplx> v(D);
      for (uint32_t g = 0; g < ngroups; ++g) {
        uint32_t lb = g;
        for (uint32_t i = 0; i < op.nfixed; ++i) {
          uint32_t p = op.fixed[i];
          lb = ((lb >> p) << (p + 1)) | (lb & ((1u << p) - 1));
        }
        lb |= op.lcmask;
        for (uint32_t j = 0; j < D; ++j) {
          check(lb + op.off[j] < tsize, "LDS index");
          v[j] = tile[lb + op.off[j]];
        }
        for (uint32_t r = 0; r < D; ++r) {
          cplx a = 0.0;
          for (uint32_t c = 0; c < D; ++c) a += mats[op.mat + r * D + c] * v[c];
          tile[lb + op.off[r]] = a;
        }
      }
    }
    for (uint32_t l = 0; l < tsize; ++l) psi[base | seg[l >> L] | (l & segmask)] = tile[l];
  }
}

}  // namespace

extern "C" {

// gates: k[i] targets at tpos[off..off+k), controls cmask[i], matrix at mats[moff..] (complex
// interleaved, 4^k entries).  state: 2^n complex interleaved, updated in place.
int emu_run(unsigned n, unsigned t, unsigned lmin, int n_gates, const uint32_t *k, const uint32_t *tpos,
            const uint64_t *cmask, const double *mats, double *state, int *n_passes, char *err, int err_len) {
  try {
    std::vector<qs::PGate> gates;
    size_t to = 0, mo = 0;
    for (int i = 0; i < n_gates; ++i) {
      qs::PGate g;
      for (uint32_t l = 0; l < k[i]; ++l) g.tpos.push_back(tpos[to++]);
      size_t d = size_t(1) << k[i];
      for (size_t e = 0; e < d * d; ++e, ++mo) g.m.push_back(cplx(mats[2 * mo], mats[2 * mo + 1]));
      g.cmask = cmask[i];
      if (g.tmask() & g.cmask) continue;
      gates.push_back(g);
    }
    gates = qs::merge_single_qubit_runs(gates);
    auto passes = qs::schedule_tiles(gates, qs::TileShape{n, t, lmin});
    size_t covered = 0;
    if (!state) {  // planning only: count passes
      *n_passes = (int)passes.size();
      return 0;
    }
    for (auto &ps : passes) {
      qs::dev::TilePass P;
      std::vector<qs::dev::TileOp> ops;
      std::vector<cplx> m;
      qs::build_tile_pass(gates, ps, std::min(t, n), P, ops, m);
      replay(reinterpret_cast<cplx *>(state), n, P, ops, m);
      covered += ps.gates.size();
    }
    check(covered == gates.size(), "every gate scheduled exactly once");
    *n_passes = (int)passes.size();
    return 0;
  } catch (const Fail &f) {
    std::strncpy(err, f.what.c_str(), err_len - 1);
    return 1;
  } catch (const std::exception &e) {
    std::strncpy(err, e.what(), err_len - 1);
    return 2;
  }
}
}
