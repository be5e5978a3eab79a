// Test-only CPU model of the tile-pass kernel (k_tile_pass in projectq_amd/csrc/kernels.hip).
//
// Runs the product planner (planner.cpp: merge, schedule, build_tile_pass) on a gate list and
// replays every tile pass on the CPU with the kernel's own index arithmetic, checking each
// global and LDS index against its bounds.  Lets the CPU test-suite check that the planner's
// passes are in bounds and that their result equals the unfused gate sequence.
#include <algorithm>
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

uint32_t slot(uint32_t x, bool c64) { return c64 ? qs::dev::lds_slot_c64(x) : qs::dev::lds_slot_c128(x); }

void apply_small(std::vector<cplx> &v, uint32_t code, uint32_t rcmask, const cplx *m, unsigned rb) {
  const uint32_t kind = (code >> 8) & 15u, b0 = code & 15u, b1 = (code >> 4) & 15u, cb = code >> 12;
  // the kernels read the register control from the code; rcmask must agree with it
  check(cb <= rb && (cb == 0 ? rcmask == 0 : rcmask == (1u << (cb - 1))), "register control encoding");
  const bool one_qubit = kind < 4 || kind == qs::dev::kOp1Real || kind == qs::dev::kOp1Z;
  check(cb == 0 || (cb - 1 != b0 && (one_qubit || cb - 1 != b1)), "control on a target");
  const uint32_t rc = cb ? 1u << (cb - 1) : 0u;
  const uint32_t NR = 1u << rb;
  check(b0 < rb && (one_qubit || (b1 < rb && b0 < b1)), "register op bits");
  if (one_qubit) {
    for (uint32_t i = 0; i < NR; ++i) {
      if ((i & rc) != rc) continue;
      if (kind == qs::dev::kOp1Z) {
        if ((i >> b0) & 1) v[i] = -v[i];
        continue;
      }
      if (kind == qs::dev::kOp1Diag) {
        v[i] *= ((i >> b0) & 1) ? m[3] : m[0];
        continue;
      }
      if ((i >> b0) & 1) continue;
      const uint32_t j = i | (1u << b0);
      const cplx x = v[i], y = v[j];
      if (kind == qs::dev::kOp1General || kind == qs::dev::kOp1Real) {
        v[i] = m[0] * x + m[1] * y;
        v[j] = m[2] * x + m[3] * y;
      } else if (kind == qs::dev::kOp1Anti) {
        v[i] = m[1] * y;
        v[j] = m[2] * x;
      } else {
        v[i] = y;
        v[j] = x;
      }
    }
    return;
  }
  for (uint32_t i = 0; i < NR; ++i) {
    if ((i & rc) != rc) continue;
    if (kind == qs::dev::kOp2Diag) {
      uint32_t s = ((i >> b0) & 1) | (((i >> b1) & 1) << 1);
      v[i] *= m[s * 5];
      continue;
    }
    if ((i >> b0) & 1 || (i >> b1) & 1) continue;
    const uint32_t idx[4] = {i, i | (1u << b0), i | (1u << b1), i | (1u << b0) | (1u << b1)};
    cplx x[4], y[4];
    for (int c = 0; c < 4; ++c) x[c] = v[idx[c]];
    for (int r = 0; r < 4; ++r) {
      y[r] = 0.0;
      for (int c = 0; c < 4; ++c) y[r] += m[r * 4 + c] * x[c];
    }
    for (int r = 0; r < 4; ++r) v[idx[r]] = y[r];
  }
}

void replay(cplx *psi, unsigned n, const qs::TileProgram &prog, unsigned rb, bool c64) {
  const qs::dev::TilePass &P = prog.P;
  const uint32_t t = P.t, L = P.L, h = P.h;
  const uint64_t N = uint64_t(1) << n;
  const uint32_t tsize = 1u << t, segmask = (1u << L) - 1;
  check(t <= n && L + h == t && h <= (uint32_t)qs::dev::kTileMaxHigh, "tile shape");
  check(P.n_stages == prog.stages.size(), "stage count");
  std::vector<uint64_t> seg(1u << h);
  for (uint32_t s = 0; s < (1u << h); ++s) {
    uint64_t o = 0;
    for (uint32_t b = 0; b < h; ++b)
      if ((s >> b) & 1) o |= uint64_t(1) << P.hpos[b];
    seg[s] = o;
  }
  const uint64_t n_tiles = N >> t;
  {
    uint32_t seen = 0;
    for (uint32_t b = 0; b < t; ++b) {
      check(P.sperm[b] < t && !((seen >> P.sperm[b]) & 1), "store relabelling is a permutation of the tile bits");
      seen |= 1u << P.sperm[b];
    }
  }
  std::vector<cplx> tile(tsize);
  auto lds = [&](uint32_t x) -> cplx & {
    check(x < tsize, "LDS index");
    uint32_t y = slot(x, c64);
    check(y < tsize, "LDS slot");
    return tile[y];
  };
  for (uint64_t tix = 0; tix < n_tiles; ++tix) {
    // the kernel's tile_base: tix above the low L bits, zeros inserted at the ascending high bits
    uint64_t base = tix << L;
    for (uint32_t b = 0; b < h; ++b) {
      check(P.hpos[b] >= L && (b == 0 || P.hpos[b] > P.hpos[b - 1]), "high bits ascend above L");
      const uint32_t p = P.hpos[b];
      base = ((base >> p) << (p + 1)) | (base & ((uint64_t(1) << p) - 1));
    }
    check(base == insert_zeros(tix, P.tiles), "tile base agrees with the tile bit list");
    check(base < N, "tile base");
    for (uint32_t l = 0; l < tsize; ++l) {
      check((l >> L) < seg.size(), "segment index");
      uint64_t gi = base | seg[l >> L] | (l & segmask);
      check(gi < N, "global index");
      lds(l) = psi[gi];
    }
    for (const auto &st : prog.stages) {
      if (st.kind == qs::dev::kStageRegisters) {
        check(st.begin <= st.end && st.end <= prog.rops.size(), "register op range");
        const uint32_t NR = 1u << rb;
        std::vector<uint32_t> roff(NR, 0);
        for (uint32_t i = 0; i < NR; ++i)
          for (uint32_t b = 0; b < rb; ++b)
            if ((i >> b) & 1) roff[i] |= 1u << st.fixed[b];
        for (uint32_t b = 1; b < rb; ++b) check(st.fixed[b - 1] < st.fixed[b] && st.fixed[b] < t, "stage bits");
        // the kernel addresses register i as slot(lb) ^ soff[i] (bytes): the planner's offsets
        // must be the swizzled register indices
        const uint32_t amp_bytes = c64 ? 8u : 16u;
        for (uint32_t i = 0; i < NR; ++i) check(st.soff[i] == slot(roff[i], c64) * amp_bytes, "stage offsets");
        auto reg_slot = [&](uint32_t lb, uint32_t i) -> cplx & {
          const uint32_t y = slot(lb, c64) ^ (st.soff[i] / amp_bytes);
          check(y < tsize && y == slot(lb | roff[i], c64), "register slot");
          return tile[y];
        };
        std::vector<cplx> v(NR);
        for (uint32_t g = 0; g < (1u << (t - rb)); ++g) {
          uint32_t lb = g;
          for (uint32_t b = 0; b < rb; ++b) {
            uint32_t p = st.fixed[b];
            lb = ((lb >> p) << (p + 1)) | (lb & ((1u << p) - 1));
          }
          for (uint32_t i = 0; i < NR; ++i) v[i] = reg_slot(lb, i);
          for (uint32_t o = st.begin; o < st.end; ++o) {
            const auto &op = prog.rops[o];
            if ((base & op.gcmask) != op.gcmask) continue;
            if ((lb & op.lcmask) != op.lcmask) continue;
            const uint32_t kd = (op.code >> 8) & 15u;
            const uint32_t width = (kd < 4 || kd == qs::dev::kOp1Real || kd == qs::dev::kOp1Z) ? 4 : 16;
            check(op.mat + width <= prog.mats.size(), "matrix offset");
            apply_small(v, op.code, op.rcmask, prog.mats.data() + op.mat, rb);
          }
          for (uint32_t i = 0; i < NR; ++i) reg_slot(lb, i) = v[i];
        }
      } else {
        check(st.begin < prog.tops.size(), "LDS op index");
        const auto &op = prog.tops[st.begin];
        if ((base & op.gcmask) != op.gcmask) continue;
        const uint32_t D = 1u << op.k;
        check(op.k >= 1 && op.k <= 5 && op.nfixed <= t, "op shape");
        check(op.mat + D * D <= prog.mats.size(), "matrix offset");
        const uint32_t ngroups = 1u << (t - op.nfixed);
        std::vector<cplx> v(D);
        for (uint32_t g = 0; g < ngroups; ++g) {
          uint32_t lb = g;
          for (uint32_t i = 0; i < op.nfixed; ++i) {
            uint32_t p = op.fixed[i];
            lb = ((lb >> p) << (p + 1)) | (lb & ((1u << p) - 1));
          }
          lb |= op.lcmask;
          for (uint32_t j = 0; j < D; ++j) v[j] = lds(lb + op.off[j]);
          for (uint32_t r = 0; r < D; ++r) {
            cplx a = 0.0;
            for (uint32_t c = 0; c < D; ++c) a += prog.mats[op.mat + r * D + c] * v[c];
            lds(lb + op.off[r]) = a;
          }
        }
      }
    }
    // the store: stored index l takes the computed amplitude at the relabelled index
    for (uint32_t l = 0; l < tsize; ++l) {
      uint32_t src = 0;
      for (uint32_t b = 0; b < t; ++b) src |= ((l >> b) & 1u) << P.sperm[b];
      psi[base | seg[l >> L] | (l & segmask)] = lds(src);
    }
  }
}

std::vector<qs::PGate> read_gates(int n_gates, const uint32_t *k, const uint32_t *tpos, const uint64_t *cmask,
                                   const double *mats) {
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
  return gates;
}

// The gate list the engine's window planners see: CNOTs rewritten as H, CZ, H where the engine
// does it (emu_set_cz), then single-qubit runs merged.
bool g_cz = false;
std::vector<qs::PGate> plan_gates(int n_gates, const uint32_t *k, const uint32_t *tpos, const uint64_t *cmask,
                                  const double *mats) {
  auto g = read_gates(n_gates, k, tpos, cmask, mats);
  return qs::merge_single_qubit_runs(g_cz ? qs::cnot_to_cz(g) : g);
}

}  // namespace

extern "C" {

void emu_set_cz(int on) { g_cz = on != 0; }

// The LDS swizzle the kernels use (layout.hpp), for the CPU tests.
uint32_t emu_slot(uint32_t x, int c64) { return slot(x, c64 != 0); }

// Planning only (no state): passes, register stages, LDS stages and register ops of the tile
// plan of a gate list, for checks on how compact the planner's passes are.
int emu_plan_stats(unsigned n, unsigned t, unsigned lmin, unsigned rb, int c64, int n_gates, const uint32_t *k,
                   const uint32_t *tpos, const uint64_t *cmask, const double *mats, int *stats,
                   uint32_t *stage_bits, int max_stage_bits, char *err, int err_len) {
  try {
    auto gates = plan_gates(n_gates, k, tpos, cmask, mats);
    auto passes = qs::schedule_tiles(gates, qs::TileShape{n, t, lmin});
    int reg = 0, lds = 0, rops = 0;
    for (auto &ps : passes) {
      qs::TileProgram prog;
      qs::build_tile_pass(gates, ps, std::min(t, n), rb, c64 != 0, prog);
      for (const auto &st : prog.stages) (st.kind == qs::dev::kStageRegisters ? reg : lds) += 1;
      rops += (int)prog.rops.size();
    }
    if (stage_bits) {  // optional: the register bits of every register stage, rb per stage
      int w = 0;
      for (auto &ps : passes) {
        qs::TileProgram prog;
        qs::build_tile_pass(gates, ps, std::min(t, n), rb, c64 != 0, prog);
        for (const auto &st : prog.stages)
          if (st.kind == qs::dev::kStageRegisters && w + (int)rb <= max_stage_bits)
            for (unsigned b = 0; b < rb; ++b) stage_bits[w++] = st.fixed[b];
      }
    }
    stats[0] = (int)passes.size();
    stats[1] = reg;
    stats[2] = lds;
    stats[3] = rops;
    stats[4] = (int)gates.size();
    return 0;
  } catch (const std::exception &e) {
    std::strncpy(err, e.what(), err_len - 1);
    return 2;
  }
}

// gates: k[i] targets at tpos[off..off+k), controls cmask[i], matrix at mats[moff..] (complex
// interleaved, 4^k entries).  state: 2^n complex interleaved, updated in place.  With `relabel`
// the passes relabel qubits at their stores (rule: slot rule of the scheduler); final_pos[label]
// receives the position each label ends at (the state is left in that layout).
int emu_run_relabel(unsigned n, unsigned t, unsigned lmin, unsigned rb, int c64, int relabel, int rule, int n_gates,
                    const uint32_t *k, const uint32_t *tpos, const uint64_t *cmask, const double *mats, double *state,
                    int *n_passes, uint32_t *final_pos, char *err, int err_len) {
  try {
    auto gates = plan_gates(n_gates, k, tpos, cmask, mats);
    auto passes = qs::schedule_tiles(gates, qs::TileShape{n, t, lmin}, qs::slot_chooser(rule), 256, 16.0, relabel != 0);
    std::vector<uint32_t> label(n), phys(n);
    for (uint32_t p = 0; p < n; ++p) label[p] = phys[p] = p;
    size_t covered = 0;
    for (auto &ps : passes) {
      if (relabel) check(ps.phys == phys, "pass positions follow the relabelling");
      qs::TileProgram prog;
      qs::build_tile_pass(gates, ps, std::min(t, n), rb, c64 != 0, prog);
      if (state) replay(reinterpret_cast<cplx *>(state), n, prog, rb, c64 != 0);
      for (const auto &sw : prog.swaps) {
        check(sw.first < n && sw.second < n, "relabelled positions");
        const uint32_t a = label[sw.first], b = label[sw.second];
        std::swap(label[sw.first], label[sw.second]);
        phys[a] = sw.second;
        phys[b] = sw.first;
      }
      covered += ps.gates.size();
    }
    check(covered == gates.size(), "every gate scheduled exactly once");
    if (final_pos)
      for (uint32_t p = 0; p < n; ++p) final_pos[p] = phys[p];
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

int emu_run(unsigned n, unsigned t, unsigned lmin, unsigned rb, int c64, int n_gates, const uint32_t *k, const uint32_t *tpos,
            const uint64_t *cmask, const double *mats, double *state, int *n_passes, char *err, int err_len) {
  try {
    auto gates = plan_gates(n_gates, k, tpos, cmask, mats);
    auto passes = qs::schedule_tiles(gates, qs::TileShape{n, t, lmin});
    size_t covered = 0;
    if (!state) {  // planning only: count passes
      *n_passes = (int)passes.size();
      return 0;
    }
    for (auto &ps : passes) {
      qs::TileProgram prog;
      qs::build_tile_pass(gates, ps, std::min(t, n), rb, c64 != 0, prog);
      replay(reinterpret_cast<cplx *>(state), n, prog, rb, c64 != 0);
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

// Register-stage ops of the plan, counted by kind (0..6): counts[kind] for ops without tile-bit
// controls, counts[7 + kind] for ops with them (lcmask != 0, the kernel's predicated forms).
extern "C" int emu_op_kinds(unsigned n, unsigned t, unsigned lmin, unsigned rb, int c64, int n_gates, const uint32_t *k,
                            const uint32_t *tpos, const uint64_t *cmask, const double *mats, int *counts, char *err,
                            int err_len) {
  try {
    auto gates = plan_gates(n_gates, k, tpos, cmask, mats);
    auto passes = qs::schedule_tiles(gates, qs::TileShape{n, t, lmin});
    for (int i = 0; i < 16; ++i) counts[i] = 0;
    for (auto &ps : passes) {
      qs::TileProgram prog;
      qs::build_tile_pass(gates, ps, std::min(t, n), rb, c64 != 0, prog);
      for (const auto &op : prog.rops) {
        const unsigned kind = (op.code >> 8) & 15u;
        if (kind < 8) counts[(op.lcmask != 0 ? 8 : 0) + kind] += 1;
      }
    }
    return 0;
  } catch (const std::exception &e) {
    std::strncpy(err, e.what(), err_len - 1);
    return 2;
  }
}

// Largest device upload of one pass program (stage, register-op, LDS-op and complex128 matrix
// arrays, each rounded up to 256 bytes as the engine's upload arena does) and the pass count.
extern "C" int emu_max_pass_upload(unsigned n, unsigned t, unsigned lmin, unsigned rb, int c64, int n_gates,
                                   const uint32_t *k, const uint32_t *tpos, const uint64_t *cmask,
                                   const double *mats, uint64_t *max_bytes, int *n_passes, char *err, int err_len) {
  try {
    auto gates = plan_gates(n_gates, k, tpos, cmask, mats);
    auto passes = qs::schedule_tiles(gates, qs::TileShape{n, t, lmin});
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    uint64_t most = 0;
    size_t covered = 0;
    for (auto &ps : passes) {
      qs::TileProgram prog;
      qs::build_tile_pass(gates, ps, std::min(t, n), rb, c64 != 0, prog);
      const uint64_t b = up(prog.stages.size() * sizeof(qs::dev::TileStage)) +
                         up(prog.rops.size() * sizeof(qs::dev::RegOp)) +
                         up(prog.tops.size() * sizeof(qs::dev::TileOp)) + up(prog.mats.size() * 16);
      most = std::max(most, b);
      covered += ps.gates.size();
    }
    check(covered == gates.size(), "every gate scheduled exactly once");
    *max_bytes = most;
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
