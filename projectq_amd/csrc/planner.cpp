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

std::vector<PPass> schedule_tiles(const std::vector<PGate> &gates, const TileShape &shape,
                                  const SlotChooser &choose, unsigned lookahead) {
  const int G = (int)gates.size();
  const unsigned n = shape.n, t = std::min(shape.t, shape.n);
  const unsigned L0 = std::min(shape.Lmin, t);
  // predecessors: the last earlier gate touching each of this gate's positions
  std::vector<std::vector<int>> preds(G);
  {
    std::vector<int> last(64, -1);
    for (int i = 0; i < G; ++i) {
      uint64_t touch = gates[i].tmask() | gates[i].cmask;
      std::set<int> ps;
      for (uint32_t p = 0; p < 64; ++p)
        if ((touch >> p) & 1) {
          if (last[p] >= 0) ps.insert(last[p]);
          last[p] = i;
        }
      preds[i].assign(ps.begin(), ps.end());
    }
  }
  std::vector<char> done(G, 0);
  std::vector<PPass> passes;
  int first = 0;
  while (first < G) {
    // forced high positions: the targets of the first unscheduled gate (always ready)
    std::set<uint32_t> high;
    for (uint32_t p : gates[first].tpos)
      if (p >= L0) high.insert(p);
    const unsigned slots = t - L0;
    // lookahead statistics over the next unscheduled gates
    std::vector<SlotCandidate> cand;
    {
      std::vector<double> score(n, 0.0), cnt(n, 0.0), firstr(n, -1.0);
      unsigned rank = 0;
      for (int i = first; i < G && rank < lookahead; ++i) {
        if (done[i]) continue;
        for (uint32_t p : gates[i].tpos) {
          if (p < L0) continue;
          score[p] += 1.0 / (1.0 + rank / 16.0);
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
    // sweep in queue order
    const int sweep_end = std::min(G, first + 8192);
    for (int i = first; i < sweep_end; ++i) {
      if (done[i]) continue;
      if ((gates[i].tmask() & ~tmask) != 0) continue;
      bool ready = true;
      for (int p : preds[i])
        if (!done[p]) {
          ready = false;
          break;
        }
      if (!ready) continue;
      done[i] = 1;
      pass.gates.push_back(i);
    }
    passes.push_back(std::move(pass));
    while (first < G && done[first]) ++first;
  }
  return passes;
}

void build_tile_pass(const std::vector<PGate> &gates, const PPass &ps, unsigned t, dev::TilePass &P,
                     std::vector<dev::TileOp> &ops, std::vector<std::complex<double>> &mats) {
  P = dev::TilePass{};
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
    P.hpos[j] = ps.high[j];
    loc[ps.high[j]] = (int)(P.L + j);
    tilepos.push_back(ps.high[j]);
  }
  std::sort(tilepos.begin(), tilepos.end());
  P.tiles.count = (uint32_t)tilepos.size();
  for (size_t i = 0; i < tilepos.size(); ++i) P.tiles.pos[i] = (uint8_t)tilepos[i];
  ops.clear();
  mats.clear();
  for (int gi : ps.gates) {
    const PGate &g = gates[gi];
    dev::TileOp op{};
    op.k = (uint32_t)g.tpos.size();
    if (op.k < 1 || op.k > 5) throw std::runtime_error("tile pass: gate width outside 1..5");
    uint32_t tb[5];
    std::vector<uint32_t> fixed;
    for (uint32_t l = 0; l < op.k; ++l) {
      if (g.tpos[l] >= 64 || loc[g.tpos[l]] < 0) throw std::runtime_error("tile planner placed a gate outside its tile");
      tb[l] = (uint32_t)loc[g.tpos[l]];
      fixed.push_back(tb[l]);
    }
    for (uint32_t p = 0; p < 64; ++p) {
      if (!((g.cmask >> p) & 1)) continue;
      if (loc[p] >= 0) {
        op.lcmask |= 1u << loc[p];
        fixed.push_back((uint32_t)loc[p]);
      } else {
        op.gcmask |= uint64_t(1) << p;
      }
    }
    std::sort(fixed.begin(), fixed.end());
    if (std::adjacent_find(fixed.begin(), fixed.end()) != fixed.end() || fixed.size() > 16 || fixed.size() > t)
      throw std::runtime_error("tile pass: repeated or too many fixed positions in one gate");
    op.nfixed = (uint32_t)fixed.size();
    for (size_t i = 0; i < fixed.size(); ++i) op.fixed[i] = fixed[i];
    for (uint32_t j = 0; j < (1u << op.k); ++j) {
      uint32_t o = 0;
      for (uint32_t l = 0; l < op.k; ++l)
        if ((j >> l) & 1) o |= 1u << tb[l];
      op.off[j] = o;
    }
    op.mat = (uint32_t)mats.size();
    mats.insert(mats.end(), g.m.begin(), g.m.end());
    ops.push_back(op);
  }
  P.n_ops = (uint32_t)ops.size();
}

}  // namespace qs
