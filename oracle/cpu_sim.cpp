// ORACLE -- test infrastructure only.
//
// A CPU restatement, in plain C++17 + OpenMP, of ProjectQ's complex128 state-vector engine
// (projectq/backends/_sim/_cppkernels/simulator.hpp, fusion.hpp and the kernel semantics of
// nointrin/kernel{1..5}.hpp).  It exports the same C ABI as the HIP library
// (include/projectq_amd.h), so the same Python binding drives both.  Only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg load it, and only as the checker.
//
// Every function names the reference lines it restates.  How each part is pinned (golden
// vectors from the compiled reference, closed forms, brute force) is listed in
// tests/test_oracle.py; nothing here is unpinned.
//
// Only complex128 is modelled: the reference has no other precision.

#include "projectq_amd.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

using cplx = std::complex<double>;
using Mat = std::vector<std::vector<cplx>>;
using Ids = std::vector<unsigned>;

thread_local std::string g_last_error;

struct CallbackError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------------------------
// Greedy gate queue: restates class Fusion (fusion.hpp:39-165).
// ---------------------------------------------------------------------------------------------
struct Queue {
  struct Item {
    Mat m;
    Ids idx;
  };
  std::set<unsigned> qubits;    // fusion.hpp:162 set_
  std::vector<Item> items;      // fusion.hpp:163 items_
  std::set<unsigned> gctrl;     // fusion.hpp:164 ctrl_set_ (controls common to every item)

  unsigned num_qubits() const { return (unsigned)qubits.size(); }

  // fusion.hpp:112-128: append controls as new top bits of the item's local index; the
  // enlarged matrix is the identity except for the block where every new control bit is 1.
  static void widen(Mat &m, Ids &idx, const Ids &extra) {
    idx.insert(idx.end(), extra.begin(), extra.end());
    size_t d = m.size(), D = d << extra.size(), off = D - d;
    Mat w(D, std::vector<cplx>(D, cplx(0.0, 0.0)));
    for (size_t i = 0; i < off; ++i) w[i][i] = 1.0;
    for (size_t r = 0; r < d; ++r)
      for (size_t c = 0; c < d; ++c) w[off + r][off + c] = m[r][c];
    m.swap(w);
  }

  // fusion.hpp:56-63 + handle_controls :130-160
  void insert(Mat m, Ids idx, const Ids &ctrl) {
    for (unsigned q : idx) qubits.insert(q);
    std::set<unsigned> stale = gctrl;  // global controls the new gate does not carry
    for (unsigned c : ctrl) {
      if (gctrl.count(c)) {
        stale.erase(c);
      } else if (!items.empty()) {
        widen(m, idx, Ids{c});
        qubits.insert(c);
      } else {
        gctrl.insert(c);
      }
    }
    if (!stale.empty()) {
      Ids extra(stale.begin(), stale.end());
      for (unsigned c : extra) {
        gctrl.erase(c);
        qubits.insert(c);
      }
      for (auto &it : items) widen(it.m, it.idx, extra);
    }
    items.push_back(Item{std::move(m), std::move(idx)});
  }

  // fusion.hpp:65-109: product of the embedded items, applied in insertion order.
  void fuse(Mat &M, Ids &order, Ids &ctrls) const {
    order.assign(qubits.begin(), qubits.end());
    size_t N = order.size(), D = size_t(1) << N;
    M.assign(D, std::vector<cplx>(D, cplx(0.0, 0.0)));
    for (size_t i = 0; i < D; ++i) M[i][i] = 1.0;
    std::vector<cplx> col(D);
    for (const auto &it : items) {
      std::vector<unsigned> bit(it.idx.size());
      for (size_t l = 0; l < it.idx.size(); ++l)
        bit[l] = unsigned(std::lower_bound(order.begin(), order.end(), it.idx[l]) - order.begin());
      size_t d = size_t(1) << it.idx.size();
      for (size_t k = 0; k < D; ++k) {
        for (size_t i = 0; i < D; ++i) col[i] = M[i][k];
        for (size_t i = 0; i < D; ++i) {
          size_t li = 0;
          for (size_t l = 0; l < bit.size(); ++l) li |= ((i >> bit[l]) & 1u) << l;
          cplx acc = 0.0;
          for (size_t j = 0; j < d; ++j) {
            size_t src = i;
            for (size_t l = 0; l < bit.size(); ++l) {
              size_t want = (j >> l) & 1u;
              src = (src & ~(size_t(1) << bit[l])) | (want << bit[l]);
            }
            acc += col[src] * it.m[li][j];
          }
          M[i][k] = acc;
        }
      }
    }
    ctrls.assign(gctrl.begin(), gctrl.end());
  }
};

// ---------------------------------------------------------------------------------------------
// The engine: restates class Simulator (simulator.hpp:37-579).
// ---------------------------------------------------------------------------------------------
struct Oracle {
  unsigned n = 0;
  std::vector<cplx> psi{cplx(1.0, 0.0)};
  std::map<unsigned, unsigned> pos;  // qubit id -> bit position
  Queue queue;
  std::mt19937 eng;
  std::uniform_real_distribution<double> uni{0.0, 1.0};

  explicit Oracle(uint32_t seed) : eng(seed) {}  // simulator.hpp:48-53

  size_t mask_of(const Ids &ids) {
    size_t m = 0;
    for (unsigned q : ids) m |= size_t(1) << pos[q];
    return m;
  }
  bool known(const Ids &ids) const {
    for (unsigned q : ids)
      if (!pos.count(q)) return false;
    return true;
  }

  // simulator.hpp:55-74: the new qubit becomes the top bit; upper half starts at zero.
  void allocate(unsigned id) {
    if (pos.count(id))
      throw std::runtime_error("AllocateQubit: ID already exists. Qubit IDs should be unique.");
    pos[id] = n++;
    psi.resize(size_t(1) << n, cplx(0.0, 0.0));
  }

  // kernel semantics of nointrin/kernel{1..5}.hpp: for every base index with all target bits
  // 0 and all control bits 1, v <- M v over the 2^k amplitudes base + sum_l b_l 2^{p_l}.
  void apply_matrix(const Mat &M, const Ids &p, size_t cmask) {
    size_t k = p.size(), d = size_t(1) << k, groups = psi.size() >> k;
    std::vector<size_t> off(d, 0);
    for (size_t j = 0; j < d; ++j)
      for (size_t l = 0; l < k; ++l)
        if ((j >> l) & 1u) off[j] |= size_t(1) << p[l];
    Ids sorted = p;
    std::sort(sorted.begin(), sorted.end());
#pragma omp parallel
    {
      std::vector<cplx> v(d);
#pragma omp for schedule(static)
      for (long long g = 0; g < (long long)groups; ++g) {
        size_t base = (size_t)g;
        for (unsigned b : sorted) base = ((base >> b) << (b + 1)) | (base & ((size_t(1) << b) - 1));
        if ((base & cmask) != cmask) continue;
        for (size_t j = 0; j < d; ++j) v[j] = psi[base + off[j]];
        for (size_t r = 0; r < d; ++r) {
          cplx acc = 0.0;
          for (size_t c = 0; c < d; ++c) acc += v[c] * M[r][c];
          psi[base + off[r]] = acc;
        }
      }
    }
  }

  // simulator.hpp:487-527
  void run() {
    if (queue.items.empty()) return;
    Mat M;
    Ids order, ctrls;
    queue.fuse(M, order, ctrls);
    if (order.empty() || order.size() > 5)
      throw std::invalid_argument("Gates with more than 5 qubits are not supported!");
    Ids p(order.size());
    for (size_t l = 0; l < order.size(); ++l) p[l] = pos[order[l]];
    apply_matrix(M, p, mask_of(ctrls));
    queue = Queue();
  }

  // simulator.hpp:204-222 with fusion_qubits_min_ = 4, fusion_qubits_max_ = 5 (:48-49)
  void apply_gate(const Mat &m, const Ids &ids, const Ids &ctrl) {
    Queue trial = queue;
    trial.insert(m, ids, ctrl);
    unsigned now = trial.num_qubits(), before = queue.num_qubits();
    if (now >= 4 && now <= 5) {
      queue = trial;
      run();
    } else if (now > 5 || (size_t)(now - ids.size()) > (size_t)before) {
      run();
      queue.insert(m, ids, ctrl);
    } else {
      queue = trial;
    }
  }

  // simulator.hpp:93-108
  bool is_classical(unsigned id, double tol) {
    run();
    size_t b = size_t(1) << pos[id];
    int lo = 0, hi = 0;
#pragma omp parallel for reduction(| : lo, hi) schedule(static)
    for (long long i = 0; i < (long long)psi.size(); ++i) {
      bool big = std::norm(psi[i]) > tol;
      if ((size_t)i & b) hi |= big; else lo |= big;
    }
    return (lo ^ hi) == 1;
  }

  // simulator.hpp:76-91: walk pairs (bit clear, bit set) in index order; the first entry above
  // tol decides.
  bool classical_value(unsigned id, double tol) {
    run();
    size_t b = size_t(1) << pos[id];
    for (size_t i = 0; i < psi.size(); ++i) {
      if (i & b) continue;
      if (std::norm(psi[i]) > tol) return false;
      if (std::norm(psi[i | b]) > tol) return true;
    }
    return false;
  }

  // simulator.hpp:110-143 (shrink = true): drop bit `pos[id]`, keeping the half with `value`.
  void remove_qubit(unsigned id, bool value) {
    run();
    unsigned p = pos[id];
    size_t b = size_t(1) << p;
    std::vector<cplx> out(psi.size() / 2);
    for (size_t j = 0; j < out.size(); ++j) {
      size_t src = ((j >> p) << (p + 1)) | (j & (b - 1)) | (value ? b : 0);
      out[j] = psi[src];
    }
    psi.swap(out);
    for (auto &kv : pos)
      if (kv.second > p) kv.second--;
    pos.erase(id);
    n--;
  }

  // simulator.hpp:194-202
  void deallocate(unsigned id) {
    run();
    if (!is_classical(id, 1e-12))
      throw std::runtime_error(
          "Error: Qubit has not been measured / uncomputed! There is most likely a bug in your code.");
    remove_qubit(id, classical_value(id, 1e-12));
  }

  // simulator.hpp:145-186: one uniform draw, a serial cumulative scan, then project and
  // renormalise.
  std::vector<uint8_t> measure(const Ids &ids) {
    run();
    double r = uni(eng);
    double acc = 0.0;
    size_t k = 0;
    while (acc < r && k < psi.size()) acc += std::norm(psi[k++]);
    size_t pick = k - 1;
    std::vector<uint8_t> out(ids.size());
    size_t m = 0, v = 0;
    for (size_t i = 0; i < ids.size(); ++i) {
      unsigned p = pos[ids[i]];
      out[i] = (pick >> p) & 1u;
      m |= size_t(1) << p;
      v |= size_t(out[i]) << p;
    }
    double kept = 0.0;
#pragma omp parallel for reduction(+ : kept) schedule(static)
    for (long long i = 0; i < (long long)psi.size(); ++i) {
      if (((size_t)i & m) != v) psi[i] = 0.0;
      else kept += std::norm(psi[i]);
    }
    double s = 1.0 / std::sqrt(kept);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)psi.size(); ++i) psi[i] *= s;
    return out;
  }

  // simulator.hpp:224-269 for a register function f(values) (serial, as in the reference).
  template <class F>
  void emulate(F &&f, std::vector<Ids> regs, const Ids &ctrl) {
    run();
    size_t cm = mask_of(ctrl);
    for (auto &r : regs)
      for (auto &q : r) q = pos[q];
    std::vector<cplx> out(psi.size(), cplx(0.0, 0.0));
    std::vector<int32_t> val(regs.size());
    for (size_t i = 0; i < psi.size(); ++i) {
      if ((i & cm) != cm) {
        out[i] += psi[i];
        continue;
      }
      for (size_t r = 0; r < regs.size(); ++r) {
        uint64_t x = 0;  // bits past 31 fall off when stored as int, as in the reference
        for (size_t b = 0; b < regs[r].size() && b < 64; ++b) x |= uint64_t((i >> regs[r][b]) & 1u) << b;
        val[r] = (int32_t)(uint32_t)x;
      }
      f(val);
      size_t j = i;
      for (size_t r = 0; r < regs.size(); ++r)
        for (size_t b = 0; b < regs[r].size(); ++b) {
          size_t want = size_t(((int64_t)val[r] >> std::min<size_t>(b, 63)) & 1);
          j = (j & ~(size_t(1) << regs[r][b])) | (want << regs[r][b]);
        }
      out[j] += psi[i];
    }
    psi.swap(out);
  }

  // simulator.hpp:538-550: apply a Pauli string as X/Y/Z single-qubit gates through the queue.
  void apply_term(const std::vector<std::pair<unsigned, char>> &term, const Ids &ids) {
    const cplx I(0.0, 1.0);
    const Mat X = {{0.0, 1.0}, {1.0, 0.0}};
    const Mat Y = {{0.0, -I}, {I, 0.0}};
    const Mat Z = {{1.0, 0.0}, {0.0, -1.0}};
    for (auto &op : term) {
      if (op.first >= ids.size()) throw std::invalid_argument("Pauli term index out of range");
      const Mat &g = op.second == 'X' ? X : op.second == 'Y' ? Y : Z;
      apply_gate(g, Ids{ids[op.first]}, Ids{});
    }
    run();
  }

  using Term = std::vector<std::pair<unsigned, char>>;

  // simulator.hpp:292-322
  double expectation(const std::vector<Term> &terms, const std::vector<double> &coef, const Ids &ids) {
    run();
    std::vector<cplx> keep = psi;
    double total = 0.0;
    for (size_t t = 0; t < terms.size(); ++t) {
      apply_term(terms[t], ids);
      double dot = 0.0;
#pragma omp parallel for reduction(+ : dot) schedule(static)
      for (long long i = 0; i < (long long)psi.size(); ++i) {
        dot += keep[i].real() * psi[i].real() + keep[i].imag() * psi[i].imag();
        psi[i] = keep[i];
      }
      total += coef[t] * dot;
    }
    return total;
  }

  // simulator.hpp:324-350
  void apply_operator(const std::vector<Term> &terms, const std::vector<cplx> &coef, const Ids &ids) {
    run();
    std::vector<cplx> keep = psi, acc(psi.size(), cplx(0.0, 0.0));
    for (size_t t = 0; t < terms.size(); ++t) {
      apply_term(terms[t], ids);
#pragma omp parallel for schedule(static)
      for (long long i = 0; i < (long long)psi.size(); ++i) {
        acc[i] += coef[t] * psi[i];
        psi[i] = keep[i];
      }
    }
    psi.swap(acc);
  }

  // simulator.hpp:386-438: scaled truncated Taylor series, stopping when the update norm on
  // the controlled subspace drops to 1e-12.
  void time_evolution(const std::vector<Term> &terms, const std::vector<double> &coef, double time,
                      const Ids &ids, const Ids &ctrl) {
    run();
    const cplx I(0.0, 1.0);
    double trace = 0.0, norm1 = 0.0;
    std::vector<Term> h;
    std::vector<double> hc;
    for (size_t t = 0; t < terms.size(); ++t) {
      if (terms[t].empty()) {
        trace += coef[t];
      } else {
        h.push_back(terms[t]);
        hc.push_back(coef[t]);
        norm1 += std::abs(coef[t]);
      }
    }
    unsigned steps = (unsigned)(std::abs(time) * norm1 + 1.0);
    cplx phase = std::exp(-time * I * trace / (double)steps);
    std::vector<cplx> out = psi;
    size_t cm = mask_of(ctrl);
    for (unsigned s = 0; s < steps; ++s) {
      double change = 1.0;
      for (unsigned k = 0; change > 1e-12; ++k) {
        cplx f = (-time * I) / double(steps * (k + 1));
        std::vector<cplx> keep = psi, upd(psi.size(), cplx(0.0, 0.0));
        for (size_t t = 0; t < h.size(); ++t) {
          apply_term(h[t], ids);
          for (size_t j = 0; j < psi.size(); ++j) {
            upd[j] += psi[j] * hc[t];
            psi[j] = keep[j];
          }
        }
        change = 0.0;
        for (size_t j = 0; j < psi.size(); ++j) {
          upd[j] *= f;
          psi[j] = upd[j];
          if ((j & cm) == cm) {
            out[j] += upd[j];
            change += std::norm(upd[j]);
          }
        }
        change = std::sqrt(change);
      }
      for (size_t j = 0; j < psi.size(); ++j) {
        if ((j & cm) == cm) out[j] *= phase;
        psi[j] = out[j];
      }
    }
  }

  // simulator.hpp:352-368
  double probability(const std::vector<uint8_t> &bits, const Ids &ids) {
    run();
    if (!known(ids))
      throw std::runtime_error(
          "get_probability(): Unknown qubit id. Please make sure you have called eng.flush().");
    size_t m = 0, v = 0;
    for (size_t i = 0; i < ids.size(); ++i) {
      m |= size_t(1) << pos[ids[i]];
      v |= size_t(bits[i] ? 1 : 0) << pos[ids[i]];
    }
    double p = 0.0;
#pragma omp parallel for reduction(+ : p) schedule(static)
    for (long long i = 0; i < (long long)psi.size(); ++i)
      if (((size_t)i & m) == v) p += std::norm(psi[i]);
    return p;
  }

  // simulator.hpp:370-384
  cplx amplitude(const std::vector<uint8_t> &bits, const Ids &ids) {
    run();
    size_t chk = 0, idx = 0;
    for (size_t i = 0; i < ids.size(); ++i) {
      if (!pos.count(ids[i])) break;
      chk |= size_t(1) << pos[ids[i]];
      idx |= size_t(bits[i] ? 1 : 0) << pos[ids[i]];
    }
    if (chk + 1 != psi.size())
      throw std::runtime_error(
          "The second argument to get_amplitude() must be a permutation of all allocated qubits. "
          "Please make sure you have called eng.flush().");
    return psi[idx];
  }

  // simulator.hpp:440-454
  void set_wavefunction(const double *wf, size_t len, const Ids &order) {
    run();
    if (pos.size() != order.size() || !known(order))
      throw std::runtime_error(
          "set_wavefunction(): Invalid mapping provided. Please make sure all qubits have been "
          "allocated previously (call eng.flush()).");
    if (len != (size_t(1) << order.size()))
      throw std::invalid_argument("set_wavefunction(): wavefunction size does not match 2^len(ordering)");
    for (size_t i = 0; i < order.size(); ++i) pos[order[i]] = (unsigned)i;
    for (size_t i = 0; i < len; ++i) psi[i] = cplx(wf[2 * i], wf[2 * i + 1]);
  }

  // simulator.hpp:456-485
  void collapse(const Ids &ids, const std::vector<uint8_t> &vals) {
    run();
    if (ids.size() != vals.size())
      throw std::length_error("collapse_wavefunction(): ids and values size mismatch");
    if (!known(ids))
      throw std::runtime_error(
          "collapse_wavefunction(): Unknown qubit id(s) provided. Try calling eng.flush() before "
          "invoking this function.");
    size_t m = 0, v = 0;
    for (size_t i = 0; i < ids.size(); ++i) {
      m |= size_t(1) << pos[ids[i]];
      v |= size_t(vals[i] ? 1 : 0) << pos[ids[i]];
    }
    double p = 0.0;
#pragma omp parallel for reduction(+ : p) schedule(static)
    for (long long i = 0; i < (long long)psi.size(); ++i)
      if (((size_t)i & m) == v) p += std::norm(psi[i]);
    if (p < 1e-12)
      throw std::runtime_error("collapse_wavefunction(): Invalid collapse! Probability is ~0.");
    double s = 1.0 / std::sqrt(p);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)psi.size(); ++i) {
      if (((size_t)i & m) != v) psi[i] = 0.0;
      else psi[i] *= s;
    }
  }
};

// ---------------------------------------------------------------------------------------------
// Integer functions of the math gates (projectq/libs/math/_gates.py), exact integer semantics
// with the int32 cast that pybind11 applies to the callback's result (_cppsim.cpp:33-41).  The
// three native fast paths use C++ int arithmetic (simulator.hpp:271-290), which wraps.
// ---------------------------------------------------------------------------------------------
int32_t to_i32(int64_t x) {
  if (x < INT32_MIN || x > INT32_MAX) throw std::runtime_error("emulate_math: result does not fit a C++ int");
  return (int32_t)x;
}
int32_t wrap32(int64_t x) { return (int32_t)(uint32_t)(uint64_t)x; }
int64_t bitlen(int64_t x) {  // len(bin(x)[2:]) for x >= 0
  if (x == 0) return 1;
  int64_t l = 0;
  while (x) { ++l; x >>= 1; }
  return l;
}
int64_t pyfloordiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
  return q;
}

void math_op(int op, const int64_t *prm, std::vector<int32_t> &v, const std::vector<size_t> &width) {
  switch (op) {
    case QS_MATH_ADD_CONSTANT:
      for (auto &x : v) x = wrap32((int64_t)x + (int32_t)prm[0]);
      break;
    case QS_MATH_ADD_CONSTANT_MOD_N:
      for (auto &x : v) x = wrap32((int32_t)wrap32((int64_t)x + (int32_t)prm[0]) % (int32_t)prm[1]);
      break;
    case QS_MATH_MULTIPLY_BY_CONSTANT_MOD_N:
      for (auto &x : v) x = wrap32((int32_t)wrap32((int64_t)x * (int32_t)prm[0]) % (int32_t)prm[1]);
      break;
    case QS_MATH_ADD_QUANTUM: {
      int64_t n = (int64_t)width[0];
      int64_t s = (int64_t)v[0] + v[1];
      if (bitlen(s) > n) {
        int64_t mod = int64_t(1) << n;
        s = ((s % mod) + mod) % mod;
        if (v.size() == 3) v[2] ^= 1;
      }
      v[1] = to_i32(s);
      break;
    }
    case QS_MATH_INVERSE_ADD_QUANTUM:
      if (v.size() == 3) v[2] ^= 1;
      v[1] = to_i32((int64_t)v[1] - v[0]);
      break;
    case QS_MATH_SUBTRACT_QUANTUM:
      v[1] = to_i32((int64_t)v[1] - v[0]);
      break;
    case QS_MATH_COMPARATOR:
      if (v[1] < v[0]) v[2] = (v[2] == 0) ? 1 : 0;
      break;
    case QS_MATH_MULTIPLY_QUANTUM:
      v[2] = to_i32((int64_t)v[2] + (int64_t)v[0] * v[1]);
      break;
    case QS_MATH_INVERSE_MULTIPLY_QUANTUM:
      v[2] = to_i32((int64_t)v[2] - (int64_t)v[0] * v[1]);
      break;
    case QS_MATH_DIVIDE_QUANTUM: {
      int64_t dividend = v[0], rem = v[1], divisor = v[2];
      if (divisor == 0 || divisor > dividend) {
        v[0] = to_i32(rem); v[1] = to_i32(dividend);
      } else {
        int64_t q = rem + pyfloordiv(dividend, divisor);
        v[0] = to_i32(dividend - q * divisor); v[1] = to_i32(q);
      }
      break;
    }
    case QS_MATH_INVERSE_DIVIDE_QUANTUM: {
      int64_t rem = v[0], quo = v[1], divisor = v[2];
      if (divisor == 0) {
        v[0] = to_i32(quo); v[1] = to_i32(rem);
      } else {
        v[0] = to_i32(rem + quo * divisor); v[1] = 0;
      }
      break;
    }
    default:
      throw std::invalid_argument("emulate_math: unknown opcode");
  }
}

size_t required_regs(int op) {
  switch (op) {
    case QS_MATH_ADD_CONSTANT:
    case QS_MATH_ADD_CONSTANT_MOD_N:
    case QS_MATH_MULTIPLY_BY_CONSTANT_MOD_N: return 1;
    case QS_MATH_ADD_QUANTUM:
    case QS_MATH_INVERSE_ADD_QUANTUM: return 2;  // a third (carry) register is optional
    case QS_MATH_SUBTRACT_QUANTUM: return 2;
    default: return 3;
  }
}

// ---------------------------------------------------------------------------------------------
// C ABI plumbing
// ---------------------------------------------------------------------------------------------
template <class F>
int guarded(F &&f) {
  try {
    f();
    return QS_OK;
  } catch (const CallbackError &e) {
    g_last_error = e.what();
    return QS_ERR_CALLBACK;
  } catch (const std::length_error &e) {
    g_last_error = e.what();
    return QS_ERR_VALUE;
  } catch (const std::invalid_argument &e) {
    g_last_error = e.what();
    return QS_ERR_VALUE;
  } catch (const std::bad_alloc &) {
    g_last_error = "out of host memory";
    return QS_ERR_MEMORY;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return QS_ERR_RUNTIME;
  }
}

Ids ids_of(const uint32_t *p, size_t n) { return Ids(p, p + n); }

std::vector<Oracle::Term> terms_of(const uint32_t *idx, const uint8_t *ops, const size_t *off, size_t nt) {
  std::vector<Oracle::Term> t(nt);
  for (size_t i = 0; i < nt; ++i)
    for (size_t j = off[i]; j < off[i + 1]; ++j) {
      char c = (char)ops[j];
      if (c != 'X' && c != 'Y' && c != 'Z') throw std::invalid_argument("Pauli op must be X, Y or Z");
      t[i].push_back({idx[j], c});
    }
  return t;
}

std::vector<Ids> regs_of(const uint32_t *q, const uint32_t *sz, size_t nr) {
  std::vector<Ids> r(nr);
  size_t k = 0;
  for (size_t i = 0; i < nr; ++i)
    for (uint32_t j = 0; j < sz[i]; ++j) r[i].push_back(q[k++]);
  return r;
}

}  // namespace

struct qs_simulator {
  Oracle o;
  explicit qs_simulator(uint32_t s) : o(s) {}
};

extern "C" {

__attribute__((visibility("default"))) const char *qs_last_error(void) { return g_last_error.c_str(); }
__attribute__((visibility("default"))) int qs_abi_version(void) { return 1; }

__attribute__((visibility("default"))) int qs_create(uint32_t seed, int precision, int fusion_mode, int planner,
                                                     uint32_t, qs_simulator **out) {
  return guarded([&] {
    if (precision != QS_C128) throw std::invalid_argument("the oracle models complex128 only");
    if (fusion_mode != QS_FUSION_REFERENCE || planner != QS_PLANNER_GREEDY)
      throw std::invalid_argument("the oracle models the reference fusion rule only");
    *out = new qs_simulator(seed);
  });
}
__attribute__((visibility("default"))) void qs_destroy(qs_simulator *s) { delete s; }

#define API __attribute__((visibility("default")))

API int qs_allocate_qubit(qs_simulator *s, uint32_t id) { return guarded([&] { s->o.allocate(id); }); }
API int qs_deallocate_qubit(qs_simulator *s, uint32_t id) {
  return guarded([&] {
    if (!s->o.pos.count(id)) throw std::runtime_error("deallocate_qubit: unknown qubit id");
    s->o.deallocate(id);
  });
}
API int qs_get_classical_value(qs_simulator *s, uint32_t id, double tol, int *v) {
  return guarded([&] {
    if (!s->o.pos.count(id)) throw std::runtime_error("get_classical_value: unknown qubit id");
    *v = s->o.classical_value(id, tol);
  });
}
API int qs_is_classical(qs_simulator *s, uint32_t id, double tol, int *v) {
  return guarded([&] {
    if (!s->o.pos.count(id)) throw std::runtime_error("is_classical: unknown qubit id");
    *v = s->o.is_classical(id, tol);
  });
}
API int qs_measure_qubits(qs_simulator *s, const uint32_t *ids, size_t n, uint8_t *res) {
  return guarded([&] {
    Ids v = ids_of(ids, n);
    if (!s->o.known(v)) throw std::runtime_error("measure_qubits: unknown qubit id");
    auto r = s->o.measure(v);
    std::copy(r.begin(), r.end(), res);
  });
}
API int qs_apply_controlled_gate(qs_simulator *s, const double *m, const uint32_t *ids, size_t k,
                                 const uint32_t *ctrl, size_t nc) {
  return guarded([&] {
    if (k > 20) throw std::invalid_argument("gate too large");
    size_t d = size_t(1) << k;
    Mat M(d, std::vector<cplx>(d));
    for (size_t r = 0; r < d; ++r)
      for (size_t c = 0; c < d; ++c) M[r][c] = cplx(m[2 * (r * d + c)], m[2 * (r * d + c) + 1]);
    Ids t = ids_of(ids, k), c = ids_of(ctrl, nc);
    if (!s->o.known(t) || !s->o.known(c)) throw std::runtime_error("apply_controlled_gate: unknown qubit id");
    s->o.apply_gate(M, t, c);
  });
}
API int qs_emulate_math(qs_simulator *s, qs_math_fn fn, void *user, const uint32_t *rq, const uint32_t *rs,
                        size_t nr, const uint32_t *ctrl, size_t nc) {
  return guarded([&] {
    auto regs = regs_of(rq, rs, nr);
    for (auto &r : regs)
      if (!s->o.known(r)) throw std::runtime_error("emulate_math: unknown qubit id");
    s->o.emulate(
        [&](std::vector<int32_t> &v) {
          if (fn(user, v.data(), v.size()) != 0) throw CallbackError("emulate_math: callback failed");
        },
        regs, ids_of(ctrl, nc));
  });
}
API int qs_emulate_math_op(qs_simulator *s, int op, const int64_t *prm, size_t np, const uint32_t *rq,
                           const uint32_t *rs, size_t nr, const uint32_t *ctrl, size_t nc) {
  return guarded([&] {
    if (op < 0 || op > QS_MATH_INVERSE_DIVIDE_QUANTUM) throw std::invalid_argument("emulate_math: unknown opcode");
    if (op <= QS_MATH_MULTIPLY_BY_CONSTANT_MOD_N) {
      size_t need = op == QS_MATH_ADD_CONSTANT ? 1 : 2;
      if (np < need) throw std::invalid_argument("emulate_math: missing constant parameters");
      if (op != QS_MATH_ADD_CONSTANT && (int32_t)prm[1] == 0) throw std::invalid_argument("emulate_math: N must be non-zero");
    } else if (nr < required_regs(op) || nr > 3) {
      throw std::invalid_argument("emulate_math: wrong number of registers for this function");
    }
    auto regs = regs_of(rq, rs, nr);
    for (auto &r : regs)
      if (!s->o.known(r)) throw std::runtime_error("emulate_math: unknown qubit id");
    std::vector<size_t> width(nr);
    for (size_t i = 0; i < nr; ++i) width[i] = rs[i];
    s->o.emulate([&](std::vector<int32_t> &v) { math_op(op, prm, v, width); }, regs, ids_of(ctrl, nc));
  });
}
API int qs_get_expectation_value(qs_simulator *s, const uint32_t *ti, const uint8_t *to, const size_t *off,
                                 size_t nt, const double *coef, const uint32_t *ids, size_t n, double *out) {
  return guarded([&] {
    Ids v = ids_of(ids, n);
    if (!s->o.known(v)) throw std::runtime_error("get_expectation_value: unknown qubit id");
    *out = s->o.expectation(terms_of(ti, to, off, nt), std::vector<double>(coef, coef + nt), v);
  });
}
API int qs_apply_qubit_operator(qs_simulator *s, const uint32_t *ti, const uint8_t *to, const size_t *off,
                                size_t nt, const double *coef, const uint32_t *ids, size_t n) {
  return guarded([&] {
    Ids v = ids_of(ids, n);
    if (!s->o.known(v)) throw std::runtime_error("apply_qubit_operator: unknown qubit id");
    std::vector<cplx> c(nt);
    for (size_t i = 0; i < nt; ++i) c[i] = cplx(coef[2 * i], coef[2 * i + 1]);
    s->o.apply_operator(terms_of(ti, to, off, nt), c, v);
  });
}
API int qs_emulate_time_evolution(qs_simulator *s, const uint32_t *ti, const uint8_t *to, const size_t *off,
                                  size_t nt, const double *coef, double time, const uint32_t *ids, size_t n,
                                  const uint32_t *ctrl, size_t nc) {
  return guarded([&] {
    Ids v = ids_of(ids, n), c = ids_of(ctrl, nc);
    if (!s->o.known(v) || !s->o.known(c)) throw std::runtime_error("emulate_time_evolution: unknown qubit id");
    s->o.time_evolution(terms_of(ti, to, off, nt), std::vector<double>(coef, coef + nt), time, v, c);
  });
}
API int qs_get_probability(qs_simulator *s, const uint8_t *bits, const uint32_t *ids, size_t n, double *out) {
  return guarded([&] { *out = s->o.probability(std::vector<uint8_t>(bits, bits + n), ids_of(ids, n)); });
}
API int qs_get_amplitude(qs_simulator *s, const uint8_t *bits, const uint32_t *ids, size_t n, double *out) {
  return guarded([&] {
    cplx a = s->o.amplitude(std::vector<uint8_t>(bits, bits + n), ids_of(ids, n));
    out[0] = a.real();
    out[1] = a.imag();
  });
}
API int qs_set_wavefunction(qs_simulator *s, const double *wf, size_t len, const uint32_t *ord, size_t n) {
  return guarded([&] { s->o.set_wavefunction(wf, len, ids_of(ord, n)); });
}
API int qs_collapse_wavefunction(qs_simulator *s, const uint32_t *ids, size_t n, const uint8_t *vals, size_t nv) {
  return guarded([&] { s->o.collapse(ids_of(ids, n), std::vector<uint8_t>(vals, vals + nv)); });
}
API int qs_run(qs_simulator *s) { return guarded([&] { s->o.run(); }); }
API int qs_cheat_sizes(qs_simulator *s, size_t *nq, size_t *na) {
  return guarded([&] {
    s->o.run();
    *nq = s->o.pos.size();
    *na = s->o.psi.size();
  });
}
API int qs_cheat(qs_simulator *s, uint32_t *ids, uint32_t *positions, double *state) {
  return guarded([&] {
    s->o.run();
    size_t k = 0;
    for (auto &kv : s->o.pos) {
      ids[k] = kv.first;
      positions[k] = kv.second;
      ++k;
    }
    std::memcpy(state, s->o.psi.data(), s->o.psi.size() * sizeof(cplx));
  });
}

// State files (the product's qs_save_state format, restated for cross-checks): 8-byte magic
// "QSAMDST1", uint32 version 1, uint32 precision (0 complex128, 1 complex64), uint32 qubits,
// uint32 0, uint64 amplitudes; (id, position) uint32 pairs; amplitudes.
API int qs_qubit_map(qs_simulator *s, uint32_t *ids, uint32_t *positions, size_t cap, size_t *n) {
  return guarded([&] {
    s->o.run();
    size_t k = 0;
    for (auto &kv : s->o.pos) {
      if (k < cap) {
        ids[k] = kv.first;
        positions[k] = kv.second;
      }
      ++k;
    }
    *n = k;
  });
}
API int qs_device_state(qs_simulator *, void **, uint64_t *, int *) {
  return guarded([&] { throw std::runtime_error("the oracle keeps its state in host memory"); });
}
API int qs_expectation_value_noalloc(const void *, uint32_t, int, const uint32_t *, const uint8_t *, const size_t *,
                                    size_t, const double *, const uint32_t *, size_t, double *) {
  return guarded([&] { throw std::runtime_error("the oracle keeps its state in host memory"); });
}
API int qs_set_wavefunction_device(qs_simulator *, const void *, uint64_t, int, const uint32_t *, size_t) {
  return guarded([&] { throw std::runtime_error("the oracle keeps its state in host memory"); });
}
API int qs_save_state(qs_simulator *s, const char *path) {
  return guarded([&] {
    s->o.run();
    FILE *f = std::fopen(path, "wb");
    if (!f) throw std::runtime_error("cannot open state file");
    const uint32_t nq = (uint32_t)s->o.pos.size();
    const uint64_t na = s->o.psi.size();
    const uint32_t head[4] = {1u, 0u, nq, 0u};
    bool ok = std::fwrite("QSAMDST1", 1, 8, f) == 8 && std::fwrite(head, 4, 4, f) == 4 && std::fwrite(&na, 8, 1, f) == 1;
    for (auto &kv : s->o.pos) {
      const uint32_t rec[2] = {kv.first, kv.second};
      ok = ok && std::fwrite(rec, 4, 2, f) == 2;
    }
    ok = ok && std::fwrite(s->o.psi.data(), sizeof(cplx), na, f) == na;
    std::fclose(f);
    if (!ok) throw std::runtime_error("save_state: write failed");
  });
}
API int qs_load_state(qs_simulator *s, const char *path) {
  return guarded([&] {
    s->o.run();
    FILE *f = std::fopen(path, "rb");
    if (!f) throw std::runtime_error("cannot open state file");
    char magic[8];
    uint32_t head[4];
    uint64_t na = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "QSAMDST1", 8) == 0 &&
              std::fread(head, 4, 4, f) == 4 && std::fread(&na, 8, 1, f) == 1 && head[0] == 1 && head[1] <= 1 &&
              head[2] <= 30 && na == (uint64_t(1) << head[2]);
    std::map<unsigned, unsigned> map;
    for (uint32_t i = 0; ok && i < head[2]; ++i) {
      uint32_t rec[2];
      ok = std::fread(rec, 4, 2, f) == 2 && rec[1] < head[2] && !map.count(rec[0]);
      if (ok) map[rec[0]] = rec[1];
    }
    std::vector<cplx> amps(ok ? na : 0);
    if (ok && head[1] == 0) {
      ok = std::fread(amps.data(), sizeof(cplx), na, f) == na;
    } else if (ok) {
      std::vector<float> v(2 * na);
      ok = std::fread(v.data(), 8, na, f) == na;
      for (uint64_t i = 0; ok && i < na; ++i) amps[i] = cplx(v[2 * i], v[2 * i + 1]);
    }
    ok = ok && std::fgetc(f) == EOF;  // no trailing bytes
    std::fclose(f);
    if (!ok) throw std::invalid_argument("load_state: malformed state file");
    if (s->o.pos.empty()) {
      std::vector<unsigned> by_pos(head[2]);
      for (auto &kv : map) by_pos[kv.second] = kv.first;
      for (unsigned id : by_pos) s->o.allocate(id);
    } else {
      bool same = s->o.pos.size() == map.size();
      for (auto &kv : map) same = same && s->o.pos.count(kv.first);
      if (!same) throw std::runtime_error("load_state: the simulator holds other qubits than the file");
    }
    s->o.pos = map;
    s->o.psi = amps;
  });
}

API int qs_synchronize(qs_simulator *) { return QS_OK; }
API int qs_set_profiling(qs_simulator *, int) { return QS_OK; }
API int qs_get_stats(qs_simulator *, uint64_t *passes, double *bytes, double *ms, uint64_t *timed) {
  *passes = 0;
  *bytes = 0.0;
  *ms = 0.0;
  *timed = 0;
  return QS_OK;
}
API int qs_reset_stats(qs_simulator *) { return QS_OK; }
API int qs_overlap(qs_simulator *a, qs_simulator *b, double *out) {
  return guarded([&] {
    a->o.run();
    b->o.run();
    if (a->o.pos != b->o.pos) throw std::invalid_argument("overlap: different qubits or positions");
    cplx acc = 0.0;
    for (size_t i = 0; i < a->o.psi.size(); ++i) acc += std::conj(a->o.psi[i]) * b->o.psi[i];
    out[0] = acc.real();
    out[1] = acc.imag();
  });
}
API int qs_compare_states(qs_simulator *a, qs_simulator *b, double *max_abs_diff, uint64_t *n_unequal) {
  return guarded([&] {
    a->o.run();
    b->o.run();
    if (a->o.pos != b->o.pos) throw std::invalid_argument("compare: different qubits or positions");
    double m = 0.0;
    uint64_t c = 0;
    for (size_t i = 0; i < a->o.psi.size(); ++i) {
      m = std::max(m, std::abs(a->o.psi[i] - b->o.psi[i]));
      c += a->o.psi[i] != b->o.psi[i];
    }
    *max_abs_diff = m;
    *n_unequal = c;
  });
}

// Planner controls and the MLP are GPU-side features with no reference counterpart; the oracle
// only answers for the simulator surface.
API int qs_set_tile_params(qs_simulator *, uint32_t, uint32_t, uint32_t) { return QS_OK; }
API int qs_set_slot_rule(qs_simulator *, int) { return QS_OK; }
API int qs_set_relabel(qs_simulator *, int) { return QS_OK; }
API int qs_last_plan_relabels(qs_simulator *, uint32_t *passes) {
  *passes = 0;
  return QS_OK;
}
API int qs_last_plan_rule(qs_simulator *, int *rule) {
  *rule = 0;
  return QS_OK;
}
API int qs_set_recording(qs_simulator *, int) { return QS_OK; }
API int qs_take_records(qs_simulator *, float *, float *, size_t, size_t *rows) {
  *rows = 0;
  return QS_OK;
}
API int qs_set_planner_model(qs_simulator *, const float *, size_t) { return QS_OK; }
API int qs_last_plan(qs_simulator *, int *kind, uint32_t *t, uint32_t *rb, uint32_t *low, uint32_t *p, uint32_t *c,
                     double *pred) {
  *kind = -1;
  *t = *rb = *low = *p = *c = 0;
  *pred = 0.0;
  return QS_OK;
}
API size_t qs_mlp_num_weights(void) { return 0; }
API int qs_mlp_predict(const float *, const float *, size_t, float *) {
  g_last_error = "the oracle has no MLP (see oracle/mlp.py)";
  return QS_ERR_RUNTIME;
}
API int qs_mlp_loss_grad(const float *, const float *, const float *, size_t, float *, float *) {
  g_last_error = "the oracle has no MLP (see oracle/mlp.py)";
  return QS_ERR_RUNTIME;
}
API int qs_mlp_train(const float *, const float *, size_t, int, int, float, uint32_t, const float *, float *, float *) {
  g_last_error = "the oracle has no MLP (see oracle/mlp.py)";
  return QS_ERR_RUNTIME;
}

// Batched circuits, restated one circuit at a time: gates applied unfused in order, then the
// expectation value of simulator.hpp:292-322.
struct qs_batch {
  std::vector<double> energies;
};

API int qs_batch_create(int precision, size_t nc, const uint32_t *nq, const size_t *goff, const uint32_t *gk,
                        const uint32_t *gnc, const uint32_t *gq, const double *mats, const size_t *toff,
                        const size_t *ooff, const uint32_t *oq, const uint8_t *ok, const double *co, int,
                        qs_batch **out) {
  return guarded([&] {
    if (precision != QS_C128) throw std::invalid_argument("the oracle models complex128 only");
    auto *b = new qs_batch();
    size_t qi = 0, mi = 0;
    for (size_t c = 0; c < nc; ++c) {
      Oracle o(1);
      for (unsigned q = 0; q < nq[c]; ++q) o.allocate(q);
      for (size_t g = goff[c]; g < goff[c + 1]; ++g) {
        const size_t d = size_t(1) << gk[g];
        Mat M(d, std::vector<cplx>(d));
        Ids t, ct;
        for (uint32_t l = 0; l < gk[g]; ++l) t.push_back(gq[qi++]);
        for (uint32_t l = 0; l < gnc[g]; ++l) ct.push_back(gq[qi++]);
        for (size_t r = 0; r < d; ++r)
          for (size_t k = 0; k < d; ++k, ++mi) M[r][k] = cplx(mats[2 * mi], mats[2 * mi + 1]);
        o.apply_gate(M, t, ct);
        o.run();
      }
      std::vector<Oracle::Term> terms;
      std::vector<double> coef;
      for (size_t t = toff[c]; t < toff[c + 1]; ++t) {
        Oracle::Term term;
        for (size_t k = ooff[t]; k < ooff[t + 1]; ++k) term.push_back({oq[k], (char)ok[k]});
        terms.push_back(term);
        coef.push_back(co[t]);
      }
      Ids ids;
      for (unsigned q = 0; q < nq[c]; ++q) ids.push_back(q);
      b->energies.push_back(o.expectation(terms, coef, ids));
    }
    *out = b;
  });
}
API int qs_batch_run(qs_batch *b, int, double *energies, double *ms) {
  std::copy(b->energies.begin(), b->energies.end(), energies);
  if (ms) *ms = 0.0;
  return QS_OK;
}
API int qs_batch_shape(qs_batch *, uint32_t *t, uint32_t *rb, uint32_t *low, uint32_t *launches) {
  *t = *rb = *low = *launches = 0;
  return QS_OK;
}
API int qs_batch_relabel(qs_batch *, int *relabel) {
  *relabel = 0;
  return QS_OK;
}
API void qs_batch_destroy(qs_batch *b) { delete b; }

}  // extern "C"
