// Host-side gate queue and block fusion.
//
// GateQueue restates the reference's greedy fusion (projectq/backends/_sim/_cppkernels/
// fusion.hpp:39-165): gates are collected into one block of at most 5 qubits; controls shared
// by every gate stay global (a control mask), other controls are folded into the matrices as
// extra high bits; the block's matrix is the ordered product of the embedded gates.
#pragma once

#include <algorithm>
#include <complex>
#include <cstdint>
#include <set>
#include <vector>

namespace qs {

using cplx = std::complex<double>;

struct Gate {
  std::vector<cplx> m;          // 2^k x 2^k row-major
  std::vector<unsigned> ids;    // matrix bit l <-> ids[l]
  std::vector<unsigned> ctrl;   // qubit ids that must be 1
};

// A fused block ready for the device: matrix bit l <-> qubit id qubits[l] (ascending ids).
struct Block {
  std::vector<cplx> m;
  std::vector<unsigned> qubits;
  std::vector<unsigned> ctrl;
};

class GateQueue {
 public:
  unsigned num_qubits() const { return (unsigned)qubits_.size(); }
  size_t size() const { return items_.size(); }
  bool empty() const { return items_.empty(); }
  void clear() {
    qubits_.clear();
    items_.clear();
    gctrl_.clear();
  }

  // fusion.hpp:56-63 with handle_controls :130-160
  void insert(const Gate &g) {
    Item it{g.m, g.ids};
    for (unsigned q : g.ids) qubits_.insert(q);
    std::set<unsigned> stale = gctrl_;
    for (unsigned c : g.ctrl) {
      if (gctrl_.count(c)) {
        stale.erase(c);
      } else if (!items_.empty()) {
        widen(it, {c});
        qubits_.insert(c);
      } else {
        gctrl_.insert(c);
      }
    }
    if (!stale.empty()) {
      std::vector<unsigned> extra(stale.begin(), stale.end());
      for (unsigned c : extra) {
        gctrl_.erase(c);
        qubits_.insert(c);
      }
      for (auto &old : items_) widen(old, extra);
    }
    items_.push_back(std::move(it));
  }

  // fusion.hpp:65-109: multiply the items, in insertion order, into one 2^m x 2^m matrix.
  Block fuse() const {
    Block b;
    b.qubits.assign(qubits_.begin(), qubits_.end());
    b.ctrl.assign(gctrl_.begin(), gctrl_.end());
    const size_t N = b.qubits.size(), D = size_t(1) << N;
    std::vector<cplx> M(D * D, cplx(0.0, 0.0)), next(D * D);
    for (size_t i = 0; i < D; ++i) M[i * D + i] = 1.0;
    for (const auto &it : items_) {
      const size_t k = it.ids.size(), d = size_t(1) << k;
      std::vector<unsigned> bit(k);
      for (size_t l = 0; l < k; ++l)
        bit[l] = unsigned(std::lower_bound(b.qubits.begin(), b.qubits.end(), it.ids[l]) - b.qubits.begin());
      size_t sel = 0;  // block bits this item touches
      for (unsigned x : bit) sel |= size_t(1) << x;
      for (size_t i = 0; i < D; ++i) {
        size_t li = 0;
        for (size_t l = 0; l < k; ++l) li |= ((i >> bit[l]) & 1u) << l;
        const size_t rest = i & ~sel;
        for (size_t col = 0; col < D; ++col) {
          cplx acc = 0.0;
          for (size_t j = 0; j < d; ++j) {
            size_t src = rest;
            for (size_t l = 0; l < k; ++l) src |= ((j >> l) & 1u) << bit[l];
            acc += M[src * D + col] * it.m[li * d + j];
          }
          next[i * D + col] = acc;
        }
      }
      M.swap(next);
    }
    b.m = std::move(M);
    return b;
  }

 private:
  struct Item {
    std::vector<cplx> m;
    std::vector<unsigned> ids;
  };

  // fusion.hpp:112-128
  static void widen(Item &it, const std::vector<unsigned> &extra) {
    const size_t d = size_t(1) << it.ids.size(), D = d << extra.size(), off = D - d;
    std::vector<cplx> w(D * D, cplx(0.0, 0.0));
    for (size_t i = 0; i < off; ++i) w[i * D + i] = 1.0;
    for (size_t r = 0; r < d; ++r)
      for (size_t c = 0; c < d; ++c) w[(off + r) * D + off + c] = it.m[r * d + c];
    it.m.swap(w);
    it.ids.insert(it.ids.end(), extra.begin(), extra.end());
  }

  std::set<unsigned> qubits_;
  std::vector<Item> items_;
  std::set<unsigned> gctrl_;
};

}  // namespace qs
