// Window planners: turn a queue of gates into device passes.
//
// The tile scheduler groups gates into tile passes.  A pass streams the state once; every tile
// (2^t amplitudes sharing their non-tile bits) is staged in LDS and every gate of the pass whose
// targets lie in the tile is applied there.  Gates only move past gates on disjoint qubits, so
// the order of operations on every qubit is the queued order and results equal the unfused
// sequence up to rounding.
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <functional>
#include <vector>

#include "layout.hpp"

namespace qs {

// A queued gate resolved to bit positions (matrix bit l <-> tpos[l]).
struct PGate {
  std::vector<std::complex<double>> m;
  std::vector<uint32_t> tpos;
  uint64_t cmask = 0;
  uint64_t tmask() const {
    uint64_t v = 0;
    for (uint32_t p : tpos) v |= uint64_t(1) << p;
    return v;
  }
};

// One tile pass: low tile bits 0..L-1 plus the `high` positions; gates in execution order.
// With qubit relabelling, the gates' positions are labels: `phys` maps each label to its position
// in the state during this pass (empty: the identity), and `swaps` lists the pairs of positions
// (both in the tile) whose qubits the pass exchanges when it stores its tiles.
struct PPass {
  uint32_t L = 0;
  std::vector<uint32_t> high;
  std::vector<int> gates;
  std::vector<uint32_t> phys;
  std::vector<std::pair<uint32_t, uint32_t>> swaps;
};

struct TileShape {
  unsigned n;     // qubits in the state
  unsigned t;     // tile bits
  unsigned Lmin;  // low bits every tile keeps (contiguous segment of 2^Lmin amplitudes)
};

// Description of a candidate tile choice that a cost model can score: for every high position
// the planner could add, how much queued work it unlocks.
struct SlotCandidate {
  uint32_t pos;
  double lookahead_score;  // decayed count of queued gates targeting pos
  double first_rank;       // rank of the first queued gate targeting pos (0 = next gate)
  double count;            // queued gates in the lookahead window that target pos
};

// Returns the high positions to add beyond the forced ones (at most `slots`), best first.
using SlotChooser =
    std::function<std::vector<uint32_t>(const std::vector<SlotCandidate> &cands, unsigned slots)>;

// A run of CNOTs (controlled X, any controls) on one target, with nothing else touching the target
// in between, whose previous and next gates on the target are uncontrolled 1-qubit gates,
// becomes H, controlled Z, ..., controlled Z, H: the same unitary, and after
// merge_single_qubit_runs the H's are part of those neighbours.  The tile kernels apply the
// controlled Z as a sign flip in place (kOp1Z) instead of exchanging registers, and the
// scheduler may move a diagonal gate past gates that only read its target.
std::vector<PGate> cnot_to_cz(const std::vector<PGate> &gates);

// Merge runs of uncontrolled single-qubit gates on the same position (no other gate touching
// that position in between) into one matrix product.
std::vector<PGate> merge_single_qubit_runs(const std::vector<PGate> &gates);

// Upper bound on the device bytes one tile pass uploads (stage, op and matrix records): a pass
// is closed before a gate that would take it past this budget, so every pass program fits the
// engine's upload arena whatever the gate mix.
constexpr size_t kPassUploadBudget = size_t(4) << 20;
// Worst-case upload bytes of one gate: its complex128 matrix plus an LDS op and its own stage.
inline size_t gate_upload_bytes(const PGate &g) {
  return sizeof(std::complex<double>) * (size_t(1) << (2 * g.tpos.size())) + sizeof(dev::TileOp) +
         sizeof(dev::TileStage) + sizeof(dev::RegOp);
}

// Tile scheduling: repeatedly pick a tile (forced: the targets of the first unscheduled gate;
// the other high slots from `choose`, default = highest lookahead score) and take, in queue
// order, every unscheduled gate whose predecessors are scheduled and whose targets are in it.
// With `relabel`, each pass then moves into the low positions (present in every tile) the qubits
// of its tile that the remaining gates need first, exchanging them with low qubits needed later;
// the exchange is a free relabelling at the pass's store (PPass::swaps), and later passes see
// the gates at their new positions (PPass::phys).
std::vector<PPass> schedule_tiles(const std::vector<PGate> &gates, const TileShape &shape,
                                  const SlotChooser &choose = nullptr, unsigned lookahead = 256,
                                  double decay = 16.0, bool relabel = false);

// Device description of one tile pass.  Gates are split into stages: runs of 1- and 2-qubit
// gates whose targets fit `rb` register bits become one register stage (RegOps, specialised
// by matrix shape: general, diagonal, anti-diagonal, X); wider gates become LDS stages
// (TileOps).  Controls are split into register-bit, tile-local and outside-tile masks.
struct TileProgram {
  dev::TilePass P{};
  std::vector<dev::TileStage> stages;
  std::vector<dev::RegOp> rops;
  std::vector<dev::TileOp> tops;
  std::vector<std::complex<double>> mats;
  std::vector<std::pair<uint32_t, uint32_t>> swaps;  // positions whose qubits the store exchanges
};
void build_tile_pass(const std::vector<PGate> &gates, const PPass &pass, unsigned t, unsigned rb, bool c64,
                     TileProgram &prog);

// A Pauli string applied left to right, ops given as (bit position, 'X'|'Y'|'Z'), equals
// omega * X^x Z^z:  X_q: x ^= q;  Z_q: omega *= (-1)^{x_q}, z ^= q;  Y_q = i X_q Z_q:
// omega *= i (-1)^{x_q}, x ^= q, z ^= q.  Then (P psi)_i = omega (-1)^{popcount((i^x)&z)} psi_{i^x}.
void encode_pauli(const std::vector<std::pair<uint32_t, char>> &ops, uint64_t *x, uint64_t *z,
                  std::complex<double> *omega);
// The same encoding fed one op at a time (no allocation).
struct PauliEncoder {
  uint64_t x = 0, z = 0;
  std::complex<double> omega{1.0, 0.0};
  void add(uint32_t pos, char op);
};

// ---- cost-model features of one pass (inputs of the learned planner's MLP) -----------------
constexpr int kPassFeatures = 24;
struct PassFeatures {
  float f[kPassFeatures];
};
// a tile pass (program built with `rb` register bits) on an n-qubit state
PassFeatures tile_pass_features(const TileProgram &prog, unsigned n, bool c64, unsigned rb);
// a dense-block pass (k targets, c controls, lowest target position) on an n-qubit state
PassFeatures block_pass_features(unsigned k, unsigned n_ctrl, unsigned min_target, unsigned n, bool c64);
// the regression target: log of the pass time per 2^20 amplitudes
inline double pass_time_target(double ms, unsigned n) { return std::log(ms) - (double(n) - 20.0) * std::log(2.0); }
inline double pass_time_from_target(double y, unsigned n) { return std::exp(y + (double(n) - 20.0) * std::log(2.0)); }

// Default slot choice: highest lookahead score first.
std::vector<uint32_t> choose_by_score(const std::vector<SlotCandidate> &cands, unsigned slots);

// Rules for the free high slots of a tile.  kSlotLookahead: highest decayed lookahead score (the
// heuristic tile planner's rule).  kSlotEarliest: the positions the earliest unscheduled gates
// need, in order of first need (ties: the higher score).  On depth-100 circuits at 25-32 qubits
// the second rule plans 12-24% fewer passes on mixed 1/2/3-qubit and random-layer circuits, but
// neither rule wins everywhere; the learned planner tries both.
enum SlotRule : int { kSlotLookahead = 0, kSlotEarliest = 1 };
constexpr int kSlotRules = 2;
SlotChooser slot_chooser(int rule);  // nullptr for kSlotLookahead (schedule_tiles' default)

}  // namespace qs
