// Tile scheduler (see planner.hpp).
#include "planner.hpp"

#include <algorithm>
#include <map>
#include <set>
#include <stdexcept>

namespace qs {

std::vector<PGate> merge_single_qubit_runs(const std::vector<PGate> &gates) {
  std::vector<PGate> out;
  out.reserve(gates.size());
  std::map<uint32_t, int> last;  // position -> index in `out` of the last gate touching it
  for (const PGate &g : gates) {
    const bool single = g.tpos.size() == 1 && g.cmask == 0;
    if (single) {
      auto it = last.find(g.tpos[0]);
      if (it != last.end()) {
        PGate &prev = out[it->second];
        if (prev.tpos.size() == 1 && prev.cmask == 0) {
          // prev <- g * prev (g acts after prev)
          std::vector<std::complex<double>> m(4);
          for (int r = 0; r < 2; ++r)
            for (int c = 0; c < 2; ++c) m[r * 2 + c] = g.m[r * 2 + 0] * prev.m[0 * 2 + c] + g.m[r * 2 + 1] * prev.m[1 * 2 + c];
          prev.m = std::move(m);
          continue;
        }
      }
    }
    out.push_back(g);
    const int idx = (int)out.size() - 1;
    uint64_t touch = g.tmask() | g.cmask;
    for (uint32_t p = 0; p < 64; ++p)
      if ((touch >> p) & 1) last[p] = idx;
  }
  return out;
}

std::vector<PGate> cnot_to_cz(const std::vector<PGate> &gates) {
  const std::complex<double> zero(0.0, 0.0), one(1.0, 0.0);
  const int n = (int)gates.size();
  auto is_cnot = [&](int i) {
    const PGate &g = gates[i];
    return g.tpos.size() == 1 && g.cmask != 0 && g.m.size() == 4 && g.m[0] == zero && g.m[1] == one &&
           g.m[2] == one && g.m[3] == zero;
  };
  auto single_on = [&](int i, uint32_t p) {
    return i >= 0 && i < n && gates[i].tpos.size() == 1 && gates[i].tpos[0] == p && gates[i].cmask == 0;
  };
  // gates touching each position, in order
  std::map<uint32_t, std::vector<int>> seq;
  for (int i = 0; i < n; ++i) {
    const uint64_t touch = gates[i].tmask() | gates[i].cmask;
    for (uint32_t p = 0; p < 64; ++p)
      if ((touch >> p) & 1) seq[p].push_back(i);
  }
  // runs of CNOTs on one target with nothing else touching it in between: X_1 .. X_k = H Z_1 .. Z_k H,
  // so a run between uncontrolled 1-qubit gates on the target takes one H before its first CNOT
  // and one after its last
  std::vector<char> h_before(n, 0), h_after(n, 0), as_cz(n, 0);
  for (auto &kv : seq) {
    const uint32_t t = kv.first;
    const std::vector<int> &s = kv.second;
    for (size_t a = 0; a < s.size();) {
      if (!is_cnot(s[a]) || gates[s[a]].tpos[0] != t) {
        ++a;
        continue;
      }
      size_t b = a;
      while (b + 1 < s.size() && is_cnot(s[b + 1]) && gates[s[b + 1]].tpos[0] == t) ++b;
      const int before = a > 0 ? s[a - 1] : -1, after = b + 1 < s.size() ? s[b + 1] : -1;
      if (single_on(before, t) && single_on(after, t)) {
        h_before[s[a]] = 1;
        h_after[s[b]] = 1;
        for (size_t k = a; k <= b; ++k) as_cz[s[k]] = 1;
      }
      a = b + 1;
    }
  }
  const double h = std::sqrt(0.5);
  const std::vector<std::complex<double>> H = {h, h, h, -h}, Z = {one, zero, zero, -one};
  std::vector<PGate> out;
  out.reserve(gates.size() + gates.size() / 2);
  for (int i = 0; i < n; ++i) {
    const PGate &g = gates[i];
    if (!as_cz[i]) {
      out.push_back(g);
      continue;
    }
    PGate hg, z;
    hg.tpos = z.tpos = g.tpos;
    hg.m = H;
    z.m = Z;
    z.cmask = g.cmask;
    if (h_before[i]) out.push_back(hg);
    out.push_back(std::move(z));
    if (h_after[i]) out.push_back(hg);
  }
  return out;
}

std::vector<uint32_t> choose_by_score(const std::vector<SlotCandidate> &cands, unsigned slots) {
  std::vector<SlotCandidate> c = cands;
  std::stable_sort(c.begin(), c.end(), [](const SlotCandidate &a, const SlotCandidate &b) {
    return a.lookahead_score > b.lookahead_score;
  });
  std::vector<uint32_t> out;
  for (auto &x : c) {
    if (out.size() >= slots || x.lookahead_score <= 0.0) break;
    out.push_back(x.pos);
  }
  return out;
}

std::vector<uint32_t> choose_earliest(const std::vector<SlotCandidate> &cands, unsigned slots) {
  std::vector<SlotCandidate> c = cands;
  std::stable_sort(c.begin(), c.end(), [](const SlotCandidate &a, const SlotCandidate &b) {
    return a.first_rank != b.first_rank ? a.first_rank < b.first_rank : a.lookahead_score > b.lookahead_score;
  });
  std::vector<uint32_t> out;
  for (auto &x : c) {
    if (out.size() >= slots) break;
    out.push_back(x.pos);
  }
  return out;
}

SlotChooser slot_chooser(int rule) {
  if (rule == kSlotLookahead) return nullptr;
  if (rule == kSlotEarliest) return choose_earliest;
  throw std::invalid_argument("unknown tile slot rule");
}

// A gate whose matrix is diagonal commutes with controls on its targets and with other diagonal
// gates there, so for dependencies its targets are read, like controls.
static bool diagonal_gate(const PGate &g) {
  const size_t d = size_t(1) << g.tpos.size();
  if (g.m.size() != d * d) return false;
  for (size_t r = 0; r < d; ++r)
    for (size_t c = 0; c < d; ++c)
      if (r != c && g.m[r * d + c] != std::complex<double>(0.0, 0.0)) return false;
  return true;
}

std::vector<PPass> schedule_tiles(const std::vector<PGate> &gates, const TileShape &shape,
                                  const SlotChooser &choose, unsigned lookahead, double decay, bool relabel) {
  const int G = (int)gates.size();
  const unsigned n = shape.n, t = std::min(shape.t, shape.n);
  const unsigned L0 = std::min(shape.Lmin, t);
  // predecessors.  A non-diagonal gate writes its target positions; controls and the targets of
  // diagonal gates are only read.  Readers of a position commute with each other, so a gate
  // depends on the last writer of each position it touches and, where it writes, on every reader
  // since that writer.
  std::vector<std::vector<int>> preds(G);
  {
    std::vector<int> writer(64, -1);
    std::vector<std::vector<int>> readers(64);
    for (int i = 0; i < G; ++i) {
      const bool diag = diagonal_gate(gates[i]);
      const uint64_t tm = diag ? 0 : gates[i].tmask(), cm = gates[i].cmask | (diag ? gates[i].tmask() : 0);
      std::set<int> ps;
      for (uint32_t p = 0; p < 64; ++p) {
        if (!(((tm | cm) >> p) & 1)) continue;
        if (writer[p] >= 0) ps.insert(writer[p]);
        if ((tm >> p) & 1) {
          ps.insert(readers[p].begin(), readers[p].end());
          readers[p].clear();
          writer[p] = i;
        } else {
          readers[p].push_back(i);
        }
      }
      preds[i].assign(ps.begin(), ps.end());
    }
  }
  std::vector<char> done(G, 0);
  std::vector<PPass> passes;
  // label (a position in `gates`) <-> position in the state; changed only by relabelling
  std::vector<uint32_t> phys(std::max(n, 1u)), label(std::max(n, 1u));
  for (uint32_t p = 0; p < phys.size(); ++p) phys[p] = label[p] = p;
  auto pos_of = [&](uint32_t lab) { return lab < phys.size() ? phys[lab] : lab; };
  int first = 0;
  while (first < G) {
    // forced high positions: the targets of the first unscheduled gate (always ready)
    std::set<uint32_t> high;
    for (uint32_t lab : gates[first].tpos)
      if (pos_of(lab) >= L0) high.insert(pos_of(lab));
    const unsigned slots = t - L0;
    // lookahead statistics over the next unscheduled gates
    std::vector<SlotCandidate> cand;
    {
      std::vector<double> score(n, 0.0), cnt(n, 0.0), firstr(n, -1.0);
      unsigned rank = 0;
      for (int i = first; i < G && rank < lookahead; ++i) {
        if (done[i]) continue;
        for (uint32_t lab : gates[i].tpos) {
          const uint32_t p = pos_of(lab);
          if (p < L0) continue;
          score[p] += 1.0 / (1.0 + rank / decay);
          cnt[p] += 1.0;
          if (firstr[p] < 0) firstr[p] = rank;
        }
        ++rank;
      }
      for (uint32_t p = L0; p < n; ++p)
        if (!high.count(p) && cnt[p] > 0) cand.push_back({p, score[p], firstr[p], cnt[p]});
    }
    if (high.size() < slots) {
      unsigned free_slots = slots - (unsigned)high.size();
      std::vector<uint32_t> pick = choose ? choose(cand, free_slots) : choose_by_score(cand, free_slots);
      for (uint32_t p : pick) {
        if (high.size() >= slots) break;
        if (p >= L0 && p < n) high.insert(p);
      }
    }
    // fill unused slots with the lowest free positions: longer contiguous segments
    for (uint32_t p = L0; p < n && high.size() < slots; ++p) high.insert(p);
    PPass pass;
    uint64_t tmask = (uint64_t(1) << L0) - 1;
    for (uint32_t p : high) tmask |= uint64_t(1) << p;
    uint32_t L = 0;
    while (L < 64 && ((tmask >> L) & 1)) ++L;
    pass.L = L;
    for (uint32_t p : high)
      if (p >= L) pass.high.push_back(p);
    // sweep in queue order, up to the upload budget of one pass
    const int sweep_end = std::min(G, first + 8192);
    size_t upload = 0;
    for (int i = first; i < sweep_end; ++i) {
      if (done[i]) continue;
      uint64_t tm = 0;
      for (uint32_t lab : gates[i].tpos) tm |= uint64_t(1) << pos_of(lab);
      if ((tm & ~tmask) != 0) continue;
      bool ready = true;
      for (int p : preds[i])
        if (!done[p]) {
          ready = false;
          break;
        }
      if (!ready) continue;
      const size_t cost = gate_upload_bytes(gates[i]);
      if (!pass.gates.empty() && upload + cost > kPassUploadBudget) break;
      upload += cost;
      done[i] = 1;
      pass.gates.push_back(i);
    }
    if (relabel) pass.phys = phys;
    while (first < G && done[first]) ++first;
    if (relabel && first < G && L0 > 0) {
      // the low positions take the tile's qubits that the remaining gates need first
      std::vector<double> need(phys.size(), 1e300);
      unsigned rank = 0;
      for (int i = first; i < G && rank < lookahead; ++i) {
        if (done[i]) continue;
        for (uint32_t lab : gates[i].tpos) need[pos_of(lab)] = std::min(need[pos_of(lab)], (double)rank);
        ++rank;
      }
      std::vector<uint32_t> cands;
      for (uint32_t p = 0; p < L0; ++p) cands.push_back(p);
      for (uint32_t p : high) cands.push_back(p);
      std::stable_sort(cands.begin(), cands.end(), [&](uint32_t a, uint32_t b) { return need[a] < need[b]; });
      const std::set<uint32_t> want(cands.begin(), cands.begin() + L0);
      std::vector<uint32_t> incoming;
      for (uint32_t p : want)
        if (p >= L0) incoming.push_back(p);
      size_t k = 0;
      for (uint32_t lo = 0; lo < L0 && k < incoming.size(); ++lo) {
        if (want.count(lo)) continue;
        const uint32_t hi = incoming[k++];
        pass.swaps.push_back({lo, hi});
        const uint32_t a = label[lo], b = label[hi];
        std::swap(label[lo], label[hi]);
        phys[a] = hi;
        phys[b] = lo;
      }
    }
    passes.push_back(std::move(pass));
  }
  return passes;
}

namespace {

using cplx = std::complex<double>;

uint32_t popcount32(uint32_t x) {
  uint32_t c = 0;
  for (; x; x &= x - 1) ++c;
  return c;
}

// a gate of the pass in tile-local terms
struct LocalGate {
  uint32_t k;
  uint32_t tb[5];
  uint32_t lc;  // tile-local controls
  uint64_t gc;  // controls outside the tile
  const std::vector<cplx> *m;
};

uint32_t classify1(const std::vector<cplx> &m) {
  const cplx zero(0.0, 0.0), one(1.0, 0.0);
  if (m[0] == zero && m[3] == zero) return (m[1] == one && m[2] == one) ? dev::kOp1X : dev::kOp1Anti;
  if (m[1] == zero && m[2] == zero) return (m[0] == one && m[3] == -one) ? dev::kOp1Z : dev::kOp1Diag;
  bool real = true;
  for (int i = 0; i < 4; ++i) real = real && m[i].imag() == 0.0;
  return real ? dev::kOp1Real : dev::kOp1General;
}

bool is_diag4(const std::vector<cplx> &m) {
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c)
      if (r != c && m[r * 4 + c] != cplx(0.0, 0.0)) return false;
  return true;
}

void emit_lds_op(const LocalGate &g, unsigned t, TileProgram &prog) {
  dev::TileOp op{};
  op.k = g.k;
  std::vector<uint32_t> fixed(g.tb, g.tb + g.k);
  op.lcmask = g.lc;
  op.gcmask = g.gc;
  for (uint32_t p = 0; p < 32; ++p)
    if ((g.lc >> p) & 1) fixed.push_back(p);
  std::sort(fixed.begin(), fixed.end());
  if (std::adjacent_find(fixed.begin(), fixed.end()) != fixed.end() || fixed.size() > 16 || fixed.size() > t)
    throw std::runtime_error("tile pass: repeated or too many fixed positions in one gate");
  op.nfixed = (uint32_t)fixed.size();
  for (size_t i = 0; i < fixed.size(); ++i) op.fixed[i] = fixed[i];
  for (uint32_t j = 0; j < (1u << g.k); ++j) {
    uint32_t o = 0;
    for (uint32_t l = 0; l < g.k; ++l)
      if ((j >> l) & 1) o |= 1u << g.tb[l];
    op.off[j] = o;
  }
  op.mat = (uint32_t)prog.mats.size();
  prog.mats.insert(prog.mats.end(), g.m->begin(), g.m->end());
  dev::TileStage st{};
  st.kind = dev::kStageLds;
  st.begin = (uint32_t)prog.tops.size();
  st.end = st.begin + 1;
  prog.tops.push_back(op);
  prog.stages.push_back(st);
}

// Register bits of a stage: the op targets `need` plus padding up to rb bits.  The padding is the
// choice with the fewest LDS bank conflicts (stage_conflict_degree over the lanes one LDS cycle
// serves: 32 lanes of 8-byte accesses for complex64, 16 lanes of 16-byte accesses for
// complex128), ties going to the lowest bits.
uint32_t pad_stage_bits(uint32_t need, unsigned t, unsigned rb, bool c64) {
  const uint32_t have = (uint32_t)popcount32(need);
  if (have >= rb || t <= rb) {
    uint32_t bits = need;
    for (uint32_t p = 0; p < t && (uint32_t)popcount32(bits) < rb; ++p) bits |= 1u << p;
    return bits;
  }
  std::vector<uint32_t> free_bits;
  for (uint32_t p = 0; p < t; ++p)
    if (!((need >> p) & 1)) free_bits.push_back(p);
  const uint32_t k = rb - have;
  const uint32_t lane_bits = c64 ? 5u : 4u;
  uint32_t best = 0, best_cost = ~0u;
  // enumerate k-subsets of free_bits in lexicographic order (lowest bits first)
  std::vector<uint32_t> idx(k);
  for (uint32_t i = 0; i < k; ++i) idx[i] = i;
  while (true) {
    uint32_t bits = need;
    for (uint32_t i : idx) bits |= 1u << free_bits[i];
    const uint32_t cost = dev::stage_conflict_degree(bits, t, c64, lane_bits);
    if (cost < best_cost) {
      best_cost = cost;
      best = bits;
    }
    int i = (int)k - 1;
    while (i >= 0 && idx[i] == free_bits.size() - k + (uint32_t)i) --i;
    if (i < 0) break;
    ++idx[i];
    for (uint32_t j = (uint32_t)i + 1; j < k; ++j) idx[j] = idx[j - 1] + 1;
  }
  return best;
}

void emit_reg_stage(const std::vector<LocalGate> &run, uint32_t bits, unsigned t, unsigned rb, bool c64,
                    TileProgram &prog) {
  bits = pad_stage_bits(bits, t, rb, c64);
  dev::TileStage st{};
  st.kind = dev::kStageRegisters;
  int regbit[32];
  for (int &x : regbit) x = -1;
  uint32_t nb = 0;
  for (uint32_t p = 0; p < t; ++p)
    if ((bits >> p) & 1) {
      regbit[p] = (int)nb;
      st.fixed[nb++] = p;
    }
  if (nb != rb) throw std::runtime_error("tile pass: register stage width");
  dev::fill_stage_offsets(st, rb, c64);
  st.begin = (uint32_t)prog.rops.size();
  auto close_stage = [&]() {
    st.end = (uint32_t)prog.rops.size();
    if (st.end > st.begin) prog.stages.push_back(st);
    st.begin = st.end;
  };
  for (const LocalGate &g : run) {
    dev::RegOp op{};
    op.gcmask = g.gc;
    for (uint32_t p = 0; p < 32; ++p) {
      if (!((g.lc >> p) & 1)) continue;
      if (regbit[p] >= 0) op.rcmask |= 1u << regbit[p];
      else op.lcmask |= 1u << p;
    }
    if (popcount32(op.rcmask) > 1) {
      // the register kernels take at most one register-bit control: run this gate on the LDS
      // tile between the ops before it and the ops after it
      close_stage();
      emit_lds_op(g, t, prog);
      continue;
    }
    const uint32_t cb = op.rcmask ? 1u + (uint32_t)__builtin_ctz(op.rcmask) : 0u;
    op.mat = (uint32_t)prog.mats.size();
    if (g.k == 1) {
      const uint32_t b0 = (uint32_t)regbit[g.tb[0]];
      op.code = cb << 12 | classify1(*g.m) << 8 | b0;
      prog.mats.insert(prog.mats.end(), g.m->begin(), g.m->end());
    } else {
      uint32_t b0 = (uint32_t)regbit[g.tb[0]], b1 = (uint32_t)regbit[g.tb[1]];
      std::vector<cplx> m = *g.m;
      if (b0 > b1) {  // make matrix bit 0 the lower register bit: swap index bits 0 and 1
        auto sw = [](int x) { return ((x & 1) << 1) | ((x >> 1) & 1); };
        std::vector<cplx> w(16);
        for (int r = 0; r < 4; ++r)
          for (int c = 0; c < 4; ++c) w[r * 4 + c] = m[sw(r) * 4 + sw(c)];
        m.swap(w);
        std::swap(b0, b1);
      }
      op.code = cb << 12 | (is_diag4(m) ? dev::kOp2Diag : dev::kOp2General) << 8 | b1 << 4 | b0;
      prog.mats.insert(prog.mats.end(), m.begin(), m.end());
    }
    prog.rops.push_back(op);
  }
  close_stage();
}

// Stages of one pass.  Within the tile, a non-diagonal gate writes its target bits; its tile-local
// controls, and the targets of a diagonal gate, are only read (the same commutation rules as the
// tile scheduler; controls outside the tile are read by every gate of the pass and never
// conflict).  Repeatedly: the first unplaced gate is always ready; a gate of 3-5 targets becomes
// an LDS stage; otherwise its targets are the forced register bits and the rest is chosen among
// the bits the next unplaced register gates target: every completion is tried and the one under
// which the stage absorbs the most gates wins, ties going to the fewest LDS bank conflicts (if
// those gates target fewer bits than are free, the remainder is the conflict-optimal padding).
// The stage takes -- sweeping in queue order until nothing more fits -- every ready gate whose
// targets lie in its bits and which has at most one control among them.  Gates move only past
// gates they commute with, so every bit sees its writes in the queued order.
void partition_stages(const std::vector<LocalGate> &lgs, const std::vector<char> &diag, unsigned t, unsigned rb,
                      bool c64, TileProgram &prog) {
  const int n = (int)lgs.size();
  std::vector<uint32_t> tm(n), wr(n), rd(n);
  std::vector<char> lds_only(n);
  for (int i = 0; i < n; ++i) {
    for (uint32_t l = 0; l < lgs[i].k; ++l) tm[i] |= 1u << lgs[i].tb[l];
    wr[i] = diag[i] ? 0u : tm[i];
    rd[i] = lgs[i].lc | (diag[i] ? tm[i] : 0u);
    lds_only[i] = lgs[i].k >= 3 || t < rb;
  }
  std::vector<std::vector<int>> preds(n);
  {
    std::vector<int> writer(32, -1);
    std::vector<std::vector<int>> readers(32);
    for (int i = 0; i < n; ++i) {
      std::set<int> ps;
      for (uint32_t b = 0; b < 32; ++b) {
        if (!(((wr[i] | rd[i]) >> b) & 1)) continue;
        if (writer[b] >= 0) ps.insert(writer[b]);
        if ((wr[i] >> b) & 1) {
          ps.insert(readers[b].begin(), readers[b].end());
          readers[b].clear();
          writer[b] = i;
        } else {
          readers[b].push_back(i);
        }
      }
      preds[i].assign(ps.begin(), ps.end());
    }
  }
  std::vector<char> done(n, 0);
  auto ready = [&](int i) {
    for (int p : preds[i])
      if (!done[p]) return false;
    return true;
  };
  int first = 0;
  while (first < n) {
    const int f = first;
    if (lds_only[f]) {
      emit_lds_op(lgs[f], t, prog);
      done[f] = 1;
    } else {
      // the bits the next unplaced register gates target (lookahead window)
      constexpr int kWindow = 48;
      std::vector<int> window;
      uint32_t cand = 0;
      for (int i = f; i < n && (int)window.size() < kWindow; ++i) {
        if (done[i] || lds_only[i]) continue;
        window.push_back(i);
        cand |= tm[i];
      }
      cand &= ~tm[f];
      const uint32_t lane_bits = c64 ? 5u : 4u;
      const uint32_t free_slots = rb - (uint32_t)popcount32(tm[f]);
      uint32_t bits;
      if ((uint32_t)popcount32(cand) <= free_slots) {
        bits = pad_stage_bits(tm[f] | cand, t, rb, c64);
      } else {
        // gates a stage with these bits would take, simulated over the window
        std::vector<char> sim(n);
        auto absorbed = [&](uint32_t b) {
          for (int i : window) sim[i] = done[i];
          int got = 0;
          for (bool progress = true; progress;) {
            progress = false;
            for (int i : window) {
              if (sim[i] || (tm[i] & ~b) || popcount32(lgs[i].lc & b) > 1) continue;
              bool ok = true;
              for (int p : preds[i])
                if (!(p >= f ? sim[p] : done[p])) {
                  ok = false;
                  break;
                }
              if (!ok) continue;
              sim[i] = 1;
              ++got;
              progress = true;
            }
          }
          return got;
        };
        std::vector<uint32_t> cbits;
        for (uint32_t b = 0; b < t; ++b)
          if ((cand >> b) & 1) cbits.push_back(b);
        std::vector<uint32_t> idx(free_slots);
        for (uint32_t i = 0; i < free_slots; ++i) idx[i] = i;
        int best_got = -1;
        uint32_t best_conf = ~0u;
        bits = 0;
        while (true) {
          uint32_t b = tm[f];
          for (uint32_t i : idx) b |= 1u << cbits[i];
          const int got = absorbed(b);
          const uint32_t conf = dev::stage_conflict_degree(b, t, c64, lane_bits);
          if (got > best_got || (got == best_got && conf < best_conf)) {
            best_got = got;
            best_conf = conf;
            bits = b;
          }
          int i = (int)free_slots - 1;
          while (i >= 0 && idx[i] == cbits.size() - free_slots + (uint32_t)i) --i;
          if (i < 0) break;
          ++idx[i];
          for (uint32_t j = (uint32_t)i + 1; j < free_slots; ++j) idx[j] = idx[j - 1] + 1;
        }
      }
      std::vector<LocalGate> run;
      for (bool progress = true; progress;) {
        progress = false;
        for (int i = f; i < n; ++i) {
          if (done[i] || lds_only[i] || (tm[i] & ~bits) || popcount32(lgs[i].lc & bits) > 1 || !ready(i)) continue;
          done[i] = 1;
          run.push_back(lgs[i]);
          progress = true;
        }
      }
      if (run.empty()) {  // the first gate has two controls among the padded bits: LDS op
        emit_lds_op(lgs[f], t, prog);
        done[f] = 1;
      } else {
        emit_reg_stage(run, bits, t, rb, c64, prog);
      }
    }
    while (first < n && done[first]) ++first;
  }
}

}  // namespace

void build_tile_pass(const std::vector<PGate> &gates, const PPass &ps, unsigned t, unsigned rb, bool c64,
                     TileProgram &prog) {
  prog = TileProgram{};
  dev::TilePass &P = prog.P;
  P.t = t;
  P.L = ps.L;
  P.h = (uint32_t)ps.high.size();
  if (P.L + P.h != t || P.h > (uint32_t)dev::kTileMaxHigh || t > (unsigned)dev::kTileMaxBits || P.L < 1)
    throw std::runtime_error("tile planner produced an invalid tile");
  int loc[64];
  for (int &x : loc) x = -1;
  std::vector<uint32_t> tilepos;
  for (uint32_t p = 0; p < P.L; ++p) {
    loc[p] = (int)p;
    tilepos.push_back(p);
  }
  for (uint32_t j = 0; j < P.h; ++j) {
    if (ps.high[j] < P.L || (j > 0 && ps.high[j] <= ps.high[j - 1]))
      throw std::runtime_error("tile planner: high tile positions must ascend above the low bits");
    P.hpos[j] = ps.high[j];
    loc[ps.high[j]] = (int)(P.L + j);
    tilepos.push_back(ps.high[j]);
  }
  std::sort(tilepos.begin(), tilepos.end());
  P.tiles.count = (uint32_t)tilepos.size();
  for (size_t i = 0; i < tilepos.size(); ++i) P.tiles.pos[i] = (uint8_t)tilepos[i];

  // the store's relabelling: exchanging the qubits at positions a and b exchanges tile bits
  for (uint32_t b = 0; b < (uint32_t)dev::kTileMaxBits; ++b) P.sperm[b] = b;
  for (const auto &sw : ps.swaps) {
    if (sw.first >= 64 || sw.second >= 64 || loc[sw.first] < 0 || loc[sw.second] < 0 || sw.first == sw.second)
      throw std::runtime_error("tile planner: a relabelled position lies outside its tile");
    std::swap(P.sperm[loc[sw.first]], P.sperm[loc[sw.second]]);
  }
  prog.swaps = ps.swaps;
  auto at = [&](uint32_t lab) -> uint32_t { return ps.phys.empty() ? lab : (lab < ps.phys.size() ? ps.phys[lab] : 64u); };

  std::vector<LocalGate> lgs;
  std::vector<char> diag;
  for (int gi : ps.gates) {
    const PGate &g = gates[gi];
    LocalGate lg{};
    lg.k = (uint32_t)g.tpos.size();
    if (lg.k < 1 || lg.k > 5) throw std::runtime_error("tile pass: gate width outside 1..5");
    uint32_t need = 0;
    for (uint32_t l = 0; l < lg.k; ++l) {
      const uint32_t p = at(g.tpos[l]);
      if (p >= 64 || loc[p] < 0) throw std::runtime_error("tile planner placed a gate outside its tile");
      lg.tb[l] = (uint32_t)loc[p];
      need |= 1u << lg.tb[l];
    }
    if (popcount32(need) != lg.k) throw std::runtime_error("tile pass: repeated target");
    for (uint32_t lab = 0; lab < 64; ++lab) {
      if (!((g.cmask >> lab) & 1)) continue;
      const uint32_t p = at(lab);
      if (p >= 64) throw std::runtime_error("tile pass: control outside the state");
      if (loc[p] >= 0) lg.lc |= 1u << loc[p];
      else lg.gc |= uint64_t(1) << p;
    }
    lg.m = &g.m;
    lgs.push_back(lg);
    diag.push_back(diagonal_gate(g) ? 1 : 0);
  }
  partition_stages(lgs, diag, t, rb, c64, prog);
  P.n_ops = (uint32_t)(prog.rops.size() + prog.tops.size());
  P.n_stages = (uint32_t)prog.stages.size();
}

PassFeatures tile_pass_features(const TileProgram &prog, unsigned n, bool c64, unsigned rb) {
  PassFeatures F{};
  float *f = F.f;
  const dev::TilePass &P = prog.P;
  f[0] = c64 ? 1.f : 0.f;
  f[1] = n / 32.f;
  f[2] = 0.f;
  f[3] = P.t / 13.f;
  f[4] = P.L / 13.f;
  f[5] = P.h / 9.f;
  f[6] = (P.h ? P.hpos[0] : n) / 32.f;
  f[7] = rb / 5.f;
  float reg = 0, lds = 0, kinds[7] = {0, 0, 0, 0, 0, 0, 0}, rctl = 0, lctl = 0, gctl = 0, ldsflops = 0;
  for (const auto &st : prog.stages) (st.kind == dev::kStageRegisters ? reg : lds) += 1;
  for (const auto &op : prog.rops) {
    uint32_t kd = (op.code >> 8) & 15u;
    if (kd == dev::kOp1Z) kd = dev::kOp1Diag;  // a diagonal op (cheaper than the general diagonal)
    kinds[std::min<uint32_t>(kd, 6)] += 1;
    rctl += op.rcmask != 0;
    lctl += op.lcmask != 0;
    gctl += op.gcmask != 0;
  }
  for (const auto &op : prog.tops) {
    ldsflops += float(1u << (2 * op.k));
    lctl += op.lcmask != 0;
    gctl += op.gcmask != 0;
  }
  f[8] = reg / 16.f;
  f[9] = lds / 16.f;
  f[10] = kinds[dev::kOp1General] / 32.f;
  f[11] = kinds[dev::kOp1Real] / 32.f;
  f[12] = kinds[dev::kOp1Diag] / 32.f;
  f[13] = (kinds[dev::kOp1Anti] + kinds[dev::kOp1X]) / 32.f;
  f[14] = kinds[dev::kOp2General] / 32.f;
  f[15] = kinds[dev::kOp2Diag] / 32.f;
  f[16] = rctl / 32.f;
  f[17] = lctl / 32.f;
  f[18] = gctl / 32.f;
  f[19] = ldsflops / 1024.f;
  f[23] = std::log2(1.f + float(prog.rops.size() + prog.tops.size())) / 8.f;
  return F;
}

PassFeatures block_pass_features(unsigned k, unsigned n_ctrl, unsigned min_target, unsigned n, bool c64) {
  PassFeatures F{};
  float *f = F.f;
  f[0] = c64 ? 1.f : 0.f;
  f[1] = n / 32.f;
  f[2] = 1.f;
  f[20] = k / 5.f;
  f[21] = n_ctrl / 8.f;
  f[22] = min_target / 32.f;
  f[23] = 1.f / 8.f;
  return F;
}

void PauliEncoder::add(uint32_t pos, char op) {
  if (pos >= 64) throw std::invalid_argument("Pauli string position out of range");
  const uint64_t q = uint64_t(1) << pos;
  const double sgn = (x & q) ? -1.0 : 1.0;
  switch (op) {
    case 'X': x ^= q; break;
    case 'Z': omega *= sgn; z ^= q; break;
    case 'Y': omega *= std::complex<double>(0.0, sgn); x ^= q; z ^= q; break;
    default: throw std::invalid_argument("Pauli operator must be X, Y or Z");
  }
}

void encode_pauli(const std::vector<std::pair<uint32_t, char>> &ops, uint64_t *x, uint64_t *z,
                  std::complex<double> *omega) {
  PauliEncoder e;
  for (const auto &op : ops) e.add(op.first, op.second);
  *x = e.x;
  *z = e.z;
  *omega = e.omega;
}

}  // namespace qs
