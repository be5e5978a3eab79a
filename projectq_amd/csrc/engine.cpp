// Engine implementation: host logic of every simulator method, restating
// projectq/backends/_sim/_cppkernels/simulator.hpp method by method, with the data-parallel
// loops replaced by the kernels of kernels.hip.  The state never leaves HBM except through
// cheat() and get_amplitude().
#include "engine.hpp"

#include "mfma_layout.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <future>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

namespace qs {

namespace {

BitInsert make_insert(std::vector<uint32_t> positions) {
  std::sort(positions.begin(), positions.end());
  if (positions.size() > (size_t)kMaxBits) throw ValueError("too many fixed qubits in one call");
  BitInsert ins{};
  ins.count = (uint32_t)positions.size();
  for (size_t i = 0; i < positions.size(); ++i) ins.pos[i] = (uint8_t)positions[i];
  return ins;
}

uint64_t insert_zeros_host(uint64_t g, const BitInsert &ins) {
  for (uint32_t i = 0; i < ins.count; ++i) {
    const uint32_t p = ins.pos[i];
    g = ((g >> p) << (p + 1)) | (g & ((uint64_t(1) << p) - 1));
  }
  return g;
}

std::vector<uint32_t> bits_of(uint64_t mask) {
  std::vector<uint32_t> out;
  for (uint32_t p = 0; p < 64; ++p)
    if ((mask >> p) & 1) out.push_back(p);
  return out;
}

// ||M^H M - I||_F of a d x d matrix, in FP64; with `as_fp32` the entries are first rounded to
// FP32 as the complex64 kernels use them.  Bounds the relative norm change M makes to any state.
double unitarity_defect(const std::vector<cplx> &m, size_t d, bool as_fp32) {
  auto at = [&](size_t r, size_t c) {
    const cplx v = m[r * d + c];
    return as_fp32 ? cplx((double)(float)v.real(), (double)(float)v.imag()) : v;
  };
  double f = 0.0;
  for (size_t i = 0; i < d; ++i)
    for (size_t j = 0; j < d; ++j) {
      cplx acc(0.0, 0.0);
      for (size_t r = 0; r < d; ++r) acc += std::conj(at(r, i)) * at(r, j);
      f += std::norm(acc - (i == j ? cplx(1.0, 0.0) : cplx(0.0, 0.0)));
    }
  return std::sqrt(f);
}

}  // namespace

Engine::Engine(uint32_t seed, int precision, int fusion_mode, int planner, unsigned capacity_hint)
    : c64_(precision == 1), fusion_mode_(fusion_mode), planner_kind_(planner),
      cap_hint_(capacity_hint), rng_(seed) {
  if (precision != 0 && precision != 1) throw ValueError("precision must be QS_C128 or QS_C64");
  if (fusion_mode != 0 && fusion_mode != 1) throw ValueError("unknown fusion mode");
  if (planner < 0 || planner > 2) throw ValueError("unknown planner");
  if (capacity_hint > 36) throw ValueError("capacity hint above 36 qubits");
  QS_HIP(hipGetDevice(&device_));
  // tile shape: 2^t amplitudes of LDS per workgroup and rb register bits per thread in register
  // stages.  complex64: t = 12 (32 KiB, four workgroups per CU at 123 VGPRs), rb = 4;
  // complex128: t = 12 (64 KiB, two workgroups per CU), rb = 4.
  tile_tmax_ = 12u;
  tile_rb_ = 4u;
  tile_lmin_ = c64_ ? 5u : 4u;
  if (const char *env = std::getenv("PROJECTQ_AMD_MFMA")) {
    if (!std::strcmp(env, "0")) mfma_mode_ = 0;
    else if (!std::strcmp(env, "all")) mfma_mode_ = 2;
    else if (!std::strcmp(env, "1")) mfma_mode_ = 1;
    else throw ValueError("PROJECTQ_AMD_MFMA must be 0, 1 or all");
  }
  if (const char *env = std::getenv("PROJECTQ_AMD_CZ")) {
    if (!std::strcmp(env, "0")) cnot_cz_ = false;
    else if (!std::strcmp(env, "1")) cnot_cz_ = true;
    else throw ValueError("PROJECTQ_AMD_CZ must be 0 or 1");
  }
  if (const char *env = std::getenv("PROJECTQ_AMD_RELABEL")) {
    if (!std::strcmp(env, "0")) relabel_mode_ = 0;
    else if (!std::strcmp(env, "1")) relabel_mode_ = 1;
    else throw ValueError("PROJECTQ_AMD_RELABEL must be 0 or 1");
  }
  if (const char *env = std::getenv("PROJECTQ_AMD_TILE")) {
    unsigned t = 0, rb = 0;
    if (std::sscanf(env, "%u:%u", &t, &rb) != 2 || t < 6 || t > (c64_ ? 13u : 12u) || (rb != 4 && rb != 5))
      throw ValueError("PROJECTQ_AMD_TILE must be \"t:rb\" with 6 <= t <= 13 (12 for c128) and rb 4 or 5");
    tile_tmax_ = t;
    tile_rb_ = rb;
  }
  QS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  QS_HIP(hipMalloc(&d_partials_, sizeof(double) * 2 * dev::kReduceGrid));
  QS_HIP(hipMalloc(&d_result_, sizeof(double) * 4));
  QS_HIP(hipHostMalloc((void **)&h_result_, sizeof(double) * 4, hipHostMallocDefault));
  QS_HIP(hipMalloc(&d_flags_, sizeof(int) * 4));
  QS_HIP(hipHostMalloc((void **)&h_flags_, sizeof(int) * 4, hipHostMallocDefault));
  QS_HIP(hipMalloc(&d_key_, sizeof(unsigned long long)));
  QS_HIP(hipHostMalloc((void **)&h_key_, sizeof(unsigned long long), hipHostMallocDefault));
  QS_HIP(hipMalloc(&d_arena_, kArenaBytes));
  QS_HIP(hipHostMalloc((void **)&h_arena_, kArenaBytes, hipHostMallocDefault));
  // initial state |> = [1]  (simulator.hpp:48-51)
  ensure_capacity(1);
  if (c64_) {
    float one[2] = {1.0f, 0.0f};
    QS_HIP(hipMemcpy(psi_, one, sizeof(one), hipMemcpyHostToDevice));
  } else {
    double one[2] = {1.0, 0.0};
    QS_HIP(hipMemcpy(psi_, one, sizeof(one), hipMemcpyHostToDevice));
  }
}

Engine::~Engine() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto &p : ev_used_) {
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  for (auto &r : records_) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : ev_free_) (void)hipEventDestroy(e);
  (void)hipFree(psi_);
  release_scratch();
  (void)hipFree(d_partials_);
  (void)hipFree(d_result_);
  (void)hipHostFree(h_result_);
  (void)hipFree(d_flags_);
  (void)hipHostFree(h_flags_);
  (void)hipFree(d_key_);
  (void)hipHostFree(h_key_);
  (void)hipFree(d_chunks_);
  (void)hipHostFree(h_chunks_);
  (void)hipFree(d_arena_);
  (void)hipHostFree(h_arena_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

// ------------------------------------------------------------------------------ memory
void Engine::release_scratch() {
  for (auto &b : buf_) {
    (void)hipFree(b);
    b = nullptr;
  }
}

void Engine::ensure_capacity(uint64_t need) {
  if (need <= cap_) return;
  uint64_t cap = std::max<uint64_t>(need, cap_hint_ ? (uint64_t(1) << cap_hint_) : 1);
  void *fresh = nullptr;
  QS_HIP(hipMalloc(&fresh, cap * amp_bytes()));
  if (psi_) {
    QS_HIP(hipMemcpyAsync(fresh, psi_, amps() * amp_bytes(), hipMemcpyDeviceToDevice, stream_));
    sync();
    QS_HIP(hipFree(psi_));
  }
  release_scratch();  // scratch follows the new capacity lazily
  psi_ = fresh;
  cap_ = cap;
}

void *Engine::scratch(int which) {
  if (!buf_[which]) QS_HIP(hipMalloc(&buf_[which], cap_ * amp_bytes()));
  return buf_[which];
}

const void *Engine::upload(const void *host, size_t bytes) {
  size_t need = (bytes + 255) & ~size_t(255);
  if (need > kArenaBytes) throw ValueError("argument too large for the upload arena");
  if (arena_off_ + need > kArenaBytes) {
    sync();  // every earlier copy out of the pinned arena has completed
    arena_off_ = 0;
  }
  std::memcpy(h_arena_ + arena_off_, host, bytes);
  QS_HIP(hipMemcpyAsync(d_arena_ + arena_off_, h_arena_ + arena_off_, bytes, hipMemcpyHostToDevice, stream_));
  const void *d = d_arena_ + arena_off_;
  arena_off_ += need;
  return d;
}

void Engine::sync() { QS_HIP(hipStreamSynchronize(stream_)); }

double Engine::finish(int n_partials, int width, double *second) {
  dev::finish_sum(d_partials_, n_partials, width, d_result_, stream_);
  QS_HIP(hipMemcpyAsync(h_result_, d_result_, sizeof(double) * width, hipMemcpyDeviceToHost, stream_));
  sync();
  if (second) *second = h_result_[1];
  return h_result_[0];
}

void Engine::check_known(const std::vector<unsigned> &ids, const char *what) const {
  for (unsigned q : ids)
    if (!pos_.count(q)) throw RuntimeError(std::string(what) + ": Unknown qubit id.");
}

uint64_t Engine::mask_of(const std::vector<unsigned> &ids) const {
  uint64_t m = 0;
  for (unsigned q : ids) m |= uint64_t(1) << pos_.at(q);
  return m;
}

double Engine::masked_norm(uint64_t mask, uint64_t val) {
  BitInsert ins = make_insert(bits_of(mask));
  uint64_t n_sub = amps() >> ins.count;
  dispatch([&](auto tag) {
    using R = decltype(tag);
    dev::norm_masked<R>((const amp_t<R> *)psi_, n_sub, ins, val, d_partials_, stream_);
  });
  return finish(dev::kReduceGrid, 1);
}

// ------------------------------------------------------------------------------ allocation
// simulator.hpp:55-74: the new qubit takes the top bit; the old state is the lower half and
// the upper half starts at zero.  Capacity is kept, so allocation is one memset.
void Engine::allocate_qubit(unsigned id) {
  if (pos_.count(id)) throw RuntimeError("AllocateQubit: ID already exists. Qubit IDs should be unique.");
  if (n_ >= 40) throw ValueError("too many qubits");
  ensure_capacity(amps() * 2);
  QS_HIP(hipMemsetAsync((char *)psi_ + amps() * amp_bytes(), 0, amps() * amp_bytes(), stream_));
  pointer_out_ = false;
  subspace_fix(uint64_t(1) << n_, 0);  // the new qubit is |0>: the upper half is zero
  ref_pos_[id] = n_;
  pos_[id] = n_++;
}

// simulator.hpp:93-108
bool Engine::is_classical(unsigned id, double tol) {
  run();
  check_known({id}, "is_classical");
  QS_HIP(hipMemsetAsync(d_flags_, 0, sizeof(int) * 2, stream_));
  int p = (int)pos_[id];
  dispatch([&](auto tag) {
    using R = decltype(tag);
    dev::classical_flags<R>((const amp_t<R> *)psi_, amps(), p, tol, d_flags_, stream_);
  });
  QS_HIP(hipMemcpyAsync(h_flags_, d_flags_, sizeof(int) * 2, hipMemcpyDeviceToHost, stream_));
  sync();
  return (h_flags_[0] ^ h_flags_[1]) == 1;
}

// simulator.hpp:76-91: the first pair (bit clear, bit set) in index order holding an entry
// above tol decides; the bit-clear entry is looked at first.
bool Engine::get_classical_value(unsigned id, double tol) {
  canonicalize();  // the first pair in the reference's index order decides
  check_known({id}, "get_classical_value");
  *h_key_ = ~0ull;
  QS_HIP(hipMemcpyAsync(d_key_, h_key_, sizeof(unsigned long long), hipMemcpyHostToDevice, stream_));
  int p = (int)pos_[id];
  dispatch([&](auto tag) {
    using R = decltype(tag);
    dev::first_pair<R>((const amp_t<R> *)psi_, amps() / 2, p, tol, d_key_, stream_);
  });
  QS_HIP(hipMemcpyAsync(h_key_, d_key_, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_));
  sync();
  if (*h_key_ == ~0ull) return false;
  return (*h_key_ & 1) == 1;
}

// simulator.hpp:194-202 with collapse_vector(shrink) :110-143
void Engine::deallocate_qubit(unsigned id) {
  canonicalize();
  check_known({id}, "deallocate_qubit");
  if (!is_classical(id, 1e-12))
    throw RuntimeError(
        "Error: Qubit has not been measured / uncomputed! There is most likely a bug in your code.");
  const bool value = get_classical_value(id, 1e-12);
  const unsigned p = pos_[id];
  const uint64_t half = amps() / 2;
  if (p == n_ - 1) {  // top bit: the kept half is already contiguous
    if (value)
      QS_HIP(hipMemcpyAsync(psi_, (char *)psi_ + half * amp_bytes(), half * amp_bytes(),
                            hipMemcpyDeviceToDevice, stream_));
  } else {
    void *dst = scratch(0);
    dispatch([&](auto tag) {
      using R = decltype(tag);
      dev::drop_bit<R>((const amp_t<R> *)psi_, (amp_t<R> *)dst, half, (int)p, value ? 1 : 0, stream_);
    });
    std::swap(psi_, buf_[0]);
  }
  sync();
  for (auto &kv : pos_)
    if (kv.second > p) kv.second--;
  pos_.erase(id);
  const unsigned rp = ref_pos_.at(id);
  for (auto &kv : ref_pos_)
    if (kv.second > rp) kv.second--;
  ref_pos_.erase(id);
  n_--;
  pointer_out_ = false;
  const uint64_t low = (uint64_t(1) << p) - 1;
  zmask_ = (zmask_ & low) | ((zmask_ >> 1) & ~low);
  zval_ = (zval_ & low) | ((zval_ >> 1) & ~low);
  norm_changed();
}

// ------------------------------------------------------------------------------ measurement
// simulator.hpp:145-186.  One uniform draw; the serial cumulative scan of the reference is
// replaced by chunk sums (one read pass), a host prefix over the chunks and an in-chunk scan,
// which lands on the same index unless the draw is within rounding of a boundary.
std::vector<uint8_t> Engine::measure_qubits(const std::vector<unsigned> &ids) {
  canonicalize();  // the pick scans the reference's index order
  check_known(ids, "measure_qubits");
  const double rnd = uni_(rng_);
  const uint64_t N = amps();
  // the scan runs over the subspace outside which every amplitude is zero: the same amplitudes in
  // the same index order, without the zeros
  const BitInsert sub = make_insert(bits_of(zmask_));
  const uint64_t n_sub = N >> sub.count;
  uint64_t pick;
  if (!(rnd > 0.0)) {
    pick = ~uint64_t(0);  // the reference's loop never runs and `pick--` wraps
  } else {
    const int lg = (int)std::min<unsigned>(n_ - sub.count, 12);
    const uint64_t len = uint64_t(1) << lg, chunks = n_sub >> lg;
    if (chunks > chunks_cap_) {
      (void)hipFree(d_chunks_);
      (void)hipHostFree(h_chunks_);
      d_chunks_ = h_chunks_ = nullptr;
      chunks_cap_ = 0;
      QS_HIP(hipMalloc(&d_chunks_, chunks * sizeof(double)));
      QS_HIP(hipHostMalloc((void **)&h_chunks_, chunks * sizeof(double), hipHostMallocDefault));
      chunks_cap_ = chunks;
    }
    dispatch([&](auto tag) {
      using R = decltype(tag);
      dev::chunk_norms<R>((const amp_t<R> *)psi_, n_sub, sub, zval_, lg, d_chunks_, stream_);
    });
    QS_HIP(hipMemcpyAsync(h_chunks_, d_chunks_, chunks * sizeof(double), hipMemcpyDeviceToHost, stream_));
    sync();
    double run_sum = 0.0;
    uint64_t c = 0;
    for (; c < chunks; ++c) {
      if (run_sum + h_chunks_[c] >= rnd) break;
      run_sum += h_chunks_[c];
    }
    if (c == chunks) {
      pick = N - 1;  // the draw exceeds the total norm: the scan runs off the end
    } else {
      *h_key_ = ~0ull;
      QS_HIP(hipMemcpyAsync(d_key_, h_key_, sizeof(unsigned long long), hipMemcpyHostToDevice, stream_));
      dispatch([&](auto tag) {
        using R = decltype(tag);
        dev::chunk_pick<R>((const amp_t<R> *)psi_, sub, zval_, c * len, len, run_sum, rnd, d_key_, stream_);
      });
      QS_HIP(hipMemcpyAsync(h_key_, d_key_, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_));
      sync();
      const uint64_t compact = (*h_key_ == ~0ull) ? (c + 1) * len - 1 : (uint64_t)*h_key_;
      pick = insert_zeros_host(compact, sub) | zval_;
    }
  }
  std::vector<uint8_t> res(ids.size());
  uint64_t mask = 0, val = 0;
  for (size_t i = 0; i < ids.size(); ++i) {
    unsigned p = pos_[ids[i]];
    res[i] = (uint8_t)((pick >> p) & 1);
    mask |= uint64_t(1) << p;
    val |= uint64_t(res[i]) << p;
  }
  // outcome bits that contradict fixed bits select only zero amplitudes (the reference's edge
  // cases of a draw of 0 or beyond the total)
  const double kept = ((val ^ zval_) & mask & zmask_) ? 0.0 : masked_norm(mask | zmask_, val | (zval_ & ~mask));
  double scale = 1.0 / std::sqrt(kept);
  dispatch([&](auto tag) {
    using R = decltype(tag);
    dev::collapse<R>((amp_t<R> *)psi_, n_sub, sub, zval_, mask, val, scale, stream_);
  });
  subspace_fix(mask, val);
  norm_changed();
  return res;
}

// ------------------------------------------------------------------------------ gates
// simulator.hpp:204-222 (reference fusion rule) or the window queue of the planners.
void Engine::apply_controlled_gate(Gate g) {
  check_known(g.ids, "apply_controlled_gate");
  check_known(g.ctrl, "apply_controlled_gate");
  if (g.m.size() != (size_t(1) << (2 * g.ids.size())))
    throw ValueError("apply_controlled_gate: matrix size does not match the number of qubits");
  pointer_out_ = false;
  {
    // a non-diagonal gate can move amplitudes along its targets: they leave the known-zero
    // description (controls and diagonal gates keep zero amplitudes zero)
    const size_t d = size_t(1) << g.ids.size();
    bool diagonal = true;
    for (size_t r = 0; r < d && diagonal; ++r)
      for (size_t c = 0; c < d; ++c)
        if (r != c && g.m[r * d + c] != cplx(0.0, 0.0)) {
          diagonal = false;
          break;
        }
    if (!diagonal) {
      zmask_ &= ~mask_of(g.ids);
      zval_ &= zmask_;
    }
  }
  // the complex64 norm guard must know whether a non-unitary gate is among the gates of the next
  // execution, and how far their FP32 matrices may move the norm
  bool nonunitary = false;
  double drift = 0.0;
  if (c64_) {
    const size_t d = size_t(1) << g.ids.size();
    nonunitary = unitarity_defect(g.m, d, false) > 1e-9;
    drift = unitarity_defect(g.m, d, true);
  }
  auto queued = [&]() {
    queued_nonunitary_ |= nonunitary;
    queued_drift_ += drift;
  };
  if (fusion_mode_ == 1) {
    queued();
    window_.push_back(std::move(g));
    return;
  }
  GateQueue trial = queue_;
  trial.insert(g);
  const unsigned now = trial.num_qubits(), before = queue_.num_qubits();
  if (now >= 4 && now <= 5) {
    queue_ = std::move(trial);
    queued();
    run();
  } else if (now > 5 || (size_t)(now - g.ids.size()) > (size_t)before) {
    run();
    queue_.insert(g);
    queued();
  } else {
    queued();
    queue_ = std::move(trial);
  }
}

// simulator.hpp:487-527: fuse the queue and stream the state once.
void Engine::run() {
  if (fusion_mode_ == 1) {
    flush_window();
    return;
  }
  if (queue_.empty()) return;
  Block b = queue_.fuse();
  if (b.qubits.empty() || b.qubits.size() > 5) throw ValueError("Gates with more than 5 qubits are not supported!");
  before_unitary_passes();
  exec_block(b);  // throws for > 5 qubits before the queue is cleared, like the reference
  queue_.clear();
  after_passes();
}

void Engine::before_unitary_passes() {
  if (!c64_ || queued_nonunitary_ || norm_known_) return;
  norm_ref_ = masked_norm(0, 0);
  norm_known_ = true;
  drift_bound_ = 0.0;
}

void Engine::after_passes() {
  if (!c64_) return;
  const double drift = queued_drift_;
  queued_drift_ = 0.0;
  if (queued_nonunitary_) {
    queued_nonunitary_ = false;
    norm_changed();
    return;
  }
  drift_bound_ += drift;
  if (drift_bound_ < kNormDriftBudget) return;
  const double cur = masked_norm(0, 0);
  if (cur > 0.0 && norm_ref_ > 0.0) {
    dispatch([&](auto tag) {
      using R = decltype(tag);
      dev::scale_masked<R>((amp_t<R> *)psi_, amps(), 0, std::sqrt(norm_ref_ / cur), 0.0, stream_);
    });
  }
  drift_bound_ = 0.0;
}

// Greedy planner over a window of queued gates: the reference's block rule (min 4 / max 5,
// simulator.hpp:204-222) applied gate by gate; each block is one streaming pass.
std::vector<Block> Engine::greedy_blocks(const std::vector<Gate> &gates) const {
  std::vector<Block> out;
  GateQueue q;
  for (const Gate &g : gates) {
    GateQueue trial = q;
    trial.insert(g);
    const unsigned now = trial.num_qubits(), before = q.num_qubits();
    if (now >= 4 && now <= 5) {
      out.push_back(trial.fuse());
      q.clear();
    } else if (now > 5 || (size_t)(now - g.ids.size()) > (size_t)before) {
      if (!q.empty()) out.push_back(q.fuse());
      q.clear();
      q.insert(g);
    } else {
      q = std::move(trial);
    }
  }
  if (!q.empty()) out.push_back(q.fuse());
  return out;
}

std::vector<PGate> Engine::resolve(const std::vector<Gate> &gates) const {
  std::vector<PGate> pg;
  pg.reserve(gates.size());
  for (const Gate &g : gates) {
    PGate p;
    p.m = g.m;
    for (unsigned id : g.ids) p.tpos.push_back(pos_.at(id));
    for (unsigned c : g.ctrl) p.cmask |= uint64_t(1) << pos_.at(c);
    if (p.tmask() & p.cmask) continue;  // a control that is also a target: never applies
    pg.push_back(std::move(p));
  }
  // CNOTs become H, CZ, H only where every pass will run the 1-qubit kernel, whose sign flip
  // makes the CZ cheap: in windows without 2-qubit gates (the only source of 2-qubit register ops)
  bool two_qubit_gates = false;
  for (const PGate &p : pg) two_qubit_gates = two_qubit_gates || p.tpos.size() == 2;
  return merge_single_qubit_runs(cnot_cz_ && !two_qubit_gates ? cnot_to_cz(pg) : pg);
}

// Tile size: 2^t amplitudes of LDS per workgroup, at most n bits, and small enough that a state
// of n qubits still spreads over >= 256 tiles when n is large.
unsigned Engine::tile_bits(unsigned tmax) const {
  if (n_ <= 10) return n_;
  return std::min(tmax, std::max(10u, n_ - 8));
}

std::vector<TileProgram> Engine::plan_tiles(const std::vector<PGate> &pg, unsigned tmax, unsigned rb,
                                            unsigned lmin, int rule, bool relabel, unsigned *t_used) const {
  const unsigned t = tile_bits(tmax);
  *t_used = t;
  std::vector<PPass> passes = schedule_tiles(pg, TileShape{n_, t, lmin}, slot_chooser(rule), 256, 16.0, relabel);
  std::vector<TileProgram> progs(passes.size());
  for (size_t i = 0; i < passes.size(); ++i) build_tile_pass(pg, passes[i], t, rb, c64_, progs[i]);
  return progs;
}

PassFeatures Engine::block_features(const Block &b) const {
  unsigned lo = 64;
  for (unsigned q : b.qubits) lo = std::min(lo, pos_.at(q));
  return block_pass_features((unsigned)b.qubits.size(), (unsigned)b.ctrl.size(), lo, n_, c64_);
}

// Plans the whole window before any pass runs; if planning fails the window is left queued and
// the state untouched.
void Engine::flush_window() {
  if (window_.empty()) return;
  for (const Gate &g : window_)
    if (g.ids.empty() || g.ids.size() > 5) throw ValueError("Gates with more than 5 qubits are not supported!");
  std::vector<Block> blocks;
  std::vector<TileProgram> progs;
  unsigned rb = tile_rb_;
  PlanInfo info;
  plan_window(window_, &blocks, &progs, &rb, &info);
  window_.clear();
  for (const TileProgram &p : progs) info.relabels += p.swaps.empty() ? 0u : 1u;
  last_plan_ = info;
  before_unitary_passes();
  if (info.kind == 0) {
    for (const Block &b : blocks) exec_block(b);
  } else {
    for (const TileProgram &p : progs) {
      exec_tile_program(p, rb);
      apply_relabel(p.swaps);
    }
  }
  after_passes();
}

// Restores the reference's layout (ref_pos_) after relabelling passes moved qubits.  One pass
// through LDS tiles into the scratch buffer; where that buffer cannot be allocated, in place with one exchange
// of two index bits per misplaced qubit (at most n - 1, each touching half the state).
void Engine::canonicalize() {
  run();
  if (pos_ == ref_pos_) return;
  void *dst = buf_[0];
  if (!dst && hipMalloc(&dst, cap_ * amp_bytes()) != hipSuccess) {
    (void)hipGetLastError();  // clear the failed allocation's error state
    dst = nullptr;
  }
  if (dst) {
    buf_[0] = dst;
    // dst position of every source position, and back
    std::vector<unsigned> to(n_), from(n_);
    uint64_t zm = 0, zv = 0;
    for (const auto &kv : pos_) {
      const unsigned f = kv.second, t = ref_pos_.at(kv.first);
      to[f] = t;
      from[t] = f;
      zm |= ((zmask_ >> f) & 1u) << t;
      zv |= ((zval_ >> f) & 1u) << t;
    }
    // tile: the source's lowest k positions (contiguous reads) and the sources of the
    // destination's lowest k positions (contiguous writes)
    const unsigned k = std::min(n_, 6u);
    std::vector<unsigned> tsrc;
    for (unsigned p = 0; p < k; ++p) tsrc.push_back(p);
    for (unsigned b = 0; b < k; ++b)
      if (from[b] >= k) tsrc.push_back(from[b]);
    std::vector<unsigned> tdst;  // store order: the destination's lowest k first
    for (unsigned b = 0; b < k; ++b) tdst.push_back(b);
    for (unsigned p : tsrc)
      if (to[p] >= k) tdst.push_back(to[p]);
    PermTile P{};
    P.m = (uint32_t)tsrc.size();
    for (size_t r = 0; r < tsrc.size(); ++r) P.src_pos[r] = (uint8_t)tsrc[r];
    for (size_t r = 0; r < tdst.size(); ++r) {
      P.dst_pos[r] = (uint8_t)tdst[r];
      const unsigned srcpos = from[tdst[r]];
      P.st_from[r] = (uint8_t)(std::find(tsrc.begin(), tsrc.end(), srcpos) - tsrc.begin());
    }
    for (unsigned p = 0; p < n_; ++p)
      if (std::find(tsrc.begin(), tsrc.end(), p) == tsrc.end()) {
        P.outer_src[P.n_outer] = (uint8_t)p;
        P.outer_dst[P.n_outer] = (uint8_t)to[p];
        ++P.n_outer;
      }
    dispatch([&](auto tag) {
      using R = decltype(tag);
      dev::permute_tiles<R>((const amp_t<R> *)psi_, (amp_t<R> *)dst, amps(), P, stream_);
    });
    std::swap(psi_, buf_[0]);
    sync();
    pos_ = ref_pos_;
    zmask_ = zm;
    zval_ = zv;
    return;
  }
  for (unsigned p = 0; p < n_; ++p) {
    // the qubit that belongs at position p
    unsigned want = 0;
    bool found = false;
    for (const auto &kv : ref_pos_)
      if (kv.second == p) {
        want = kv.first;
        found = true;
        break;
      }
    if (!found) throw RuntimeError("qubit layout: no qubit for a position");
    const unsigned q = pos_.at(want);
    if (q == p) continue;
    dispatch([&](auto tag) {
      using R = decltype(tag);
      dev::swap_bits<R>((amp_t<R> *)psi_, amps(), (int)p, (int)q, stream_);
    });
    apply_relabel({{p, q}});
  }
  sync();
}

// A pass that relabels exchanged the qubits at these positions when it stored its tiles: the
// qubit map and the known-zero bits follow.
void Engine::apply_relabel(const std::vector<std::pair<uint32_t, uint32_t>> &swaps) {
  if (swaps.empty()) return;
  std::vector<int64_t> id_at(n_, -1);
  for (const auto &kv : pos_) id_at[kv.second] = kv.first;
  for (const auto &sw : swaps) {
    const uint32_t a = sw.first, b = sw.second;
    if (a >= n_ || b >= n_ || id_at[a] < 0 || id_at[b] < 0) throw RuntimeError("relabelling outside the state");
    pos_[(unsigned)id_at[a]] = b;
    pos_[(unsigned)id_at[b]] = a;
    std::swap(id_at[a], id_at[b]);
    auto swap_bits = [&](uint64_t &m) {
      const uint64_t x = ((m >> a) ^ (m >> b)) & 1u;
      m ^= (x << a) | (x << b);
    };
    swap_bits(zmask_);
    swap_bits(zval_);
  }
}

void Engine::plan_window(const std::vector<Gate> &gates, std::vector<Block> *blocks_out,
                         std::vector<TileProgram> *progs_out, unsigned *rb_out, PlanInfo *info) {
  *info = PlanInfo{};
  if (planner_kind_ == 0 || (planner_kind_ == 1 && !model_)) {
    // greedy max-5 blocks: the reference's rule, and the learned planner's fallback
    *blocks_out = greedy_blocks(gates);
    info->kind = 0;
    info->passes = (unsigned)blocks_out->size();
    return;
  }
  const std::vector<PGate> pg = resolve(gates);
  if (planner_kind_ == 2) {
    unsigned t;
    *progs_out = plan_tiles(pg, tile_tmax_, tile_rb_, tile_lmin_, slot_rule_, relabel_mode_ == 1, &t);
    *rb_out = tile_rb_;
    *info = PlanInfo{1, tile_tmax_, tile_rb_, tile_lmin_, (unsigned)progs_out->size(), 1, slot_rule_, 0.0};
    return;
  }
  // learned planner: candidate plans of every tile shape under every slot rule (the groupings
  // differ), built concurrently on host threads, and the greedy max-5 blocks; the cost model
  // predicts the time of every pass of every candidate and the cheapest plan runs
  struct Cand {
    unsigned tmax, rb, lmin;
    int rule;
  };
  std::vector<Cand> cands;
  const std::vector<std::array<unsigned, 3>> shapes =
      c64_ ? std::vector<std::array<unsigned, 3>>{{12, 4, 5}, {12, 4, 6}, {12, 4, 4}, {11, 4, 5}, {10, 4, 4}, {12, 5, 5}, {13, 4, 4}, {13, 4, 5}}
           : std::vector<std::array<unsigned, 3>>{{12, 4, 4}, {12, 4, 5}, {11, 4, 4}, {10, 4, 4}, {12, 5, 4}};
  for (int rule = 0; rule < kSlotRules; ++rule)
    for (const auto &s : shapes) cands.push_back(Cand{s[0], s[1], s[2], rule});
  std::vector<std::vector<TileProgram>> plans(cands.size());
  {
    std::vector<std::future<void>> work;
    for (size_t c = 0; c < cands.size(); ++c)
      work.push_back(std::async(std::launch::async, [&, c] {
        unsigned t;
        plans[c] = plan_tiles(pg, cands[c].tmax, cands[c].rb, cands[c].lmin, cands[c].rule, relabel_mode_ != 0, &t);
      }));
    for (auto &w : work) w.get();
  }
  std::vector<float> X;
  std::vector<size_t> rows_of;
  for (size_t c = 0; c < cands.size(); ++c) {
    for (const TileProgram &p : plans[c]) {
      PassFeatures f = tile_pass_features(p, n_, c64_, cands[c].rb);
      X.insert(X.end(), f.f, f.f + kPassFeatures);
    }
    rows_of.push_back(plans[c].size());
  }
  std::vector<Block> blocks = greedy_blocks(gates);
  for (const Block &b : blocks) {
    PassFeatures f = block_features(b);
    X.insert(X.end(), f.f, f.f + kPassFeatures);
  }
  rows_of.push_back(blocks.size());
  const std::vector<float> y = model_->predict(X, X.size() / kPassFeatures, stream_);
  // Every pass streams the whole state, so a plan with several times the passes of another cannot
  // win by much; such candidates are left out, which keeps a cost model extrapolating on a gate
  // mix it was not trained on from picking a plan of many cheap-looking passes.
  size_t fewest = std::max<size_t>(rows_of[0], 1);
  for (size_t n_rows : rows_of) fewest = std::min(fewest, std::max<size_t>(n_rows, 1));
  size_t r = 0, best = rows_of.size();
  double best_ms = 0.0;
  for (size_t c = 0; c < rows_of.size(); ++c) {
    double ms = 0.0;
    for (size_t i = 0; i < rows_of[c]; ++i, ++r) ms += pass_time_from_target(y[r], n_);
    if (rows_of[c] > kMaxPassRatio * fewest) continue;
    if (best == rows_of.size() || ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  info->candidates = (unsigned)rows_of.size();
  info->predicted_ms = best_ms;
  if (best == cands.size()) {
    info->kind = 0;
    info->passes = (unsigned)blocks.size();
    *blocks_out = std::move(blocks);
    return;
  }
  const Cand &c = cands[best];
  info->kind = 1;
  info->tmax = c.tmax;
  info->rb = c.rb;
  info->lmin = c.lmin;
  info->rule = c.rule;
  info->passes = (unsigned)plans[best].size();
  *rb_out = c.rb;
  *progs_out = std::move(plans[best]);
}

void Engine::begin_pass(hipEvent_t *e0, hipEvent_t *e1) {
  *e0 = *e1 = nullptr;
  if (!profiling_ && !recording_) return;
  *e0 = take_event();
  *e1 = take_event();
  QS_HIP(hipEventRecord(*e0, stream_));
}

void Engine::end_pass(hipEvent_t e0, hipEvent_t e1, const PassFeatures *f) {
  if (!e0) return;
  QS_HIP(hipEventRecord(e1, stream_));
  if (recording_ && f) {
    records_.push_back({*f, e0, e1});  // timed through take_records, not through the stats
  } else {
    ev_used_.push_back({e0, e1});
  }
}

void Engine::exec_tile_program(const TileProgram &prog, unsigned rb) {
  if (prog.stages.empty()) return;
  const unsigned t = prog.P.t;
  const uint64_t n_tiles = amps() >> t;
  auto up = [&](const auto &v) -> const void * {
    return v.empty() ? nullptr : upload(v.data(), v.size() * sizeof(v[0]));
  };
  const dev::TileStage *d_st = (const dev::TileStage *)up(prog.stages);
  const dev::RegOp *d_rops = (const dev::RegOp *)up(prog.rops);
  const dev::TileOp *d_tops = (const dev::TileOp *)up(prog.tops);
  bool two_qubit_ops = false;
  for (const dev::RegOp &op : prog.rops) {
    const uint32_t kind = (op.code >> 8) & 15u;
    two_qubit_ops = two_qubit_ops || kind == dev::kOp2General || kind == dev::kOp2Diag;
  }
  PassFeatures f{};
  if (recording_) f = tile_pass_features(prog, n_, c64_, rb);
  dispatch([&](auto tag) {
    using R = decltype(tag);
    std::vector<amp_t<R>> m(prog.mats.size());
    for (size_t i = 0; i < m.size(); ++i) {
      m[i].x = (R)prog.mats[i].real();
      m[i].y = (R)prog.mats[i].imag();
    }
    const amp_t<R> *dm = (const amp_t<R> *)up(m);
    hipEvent_t e0, e1;
    begin_pass(&e0, &e1);
    dev::tile_pass<R>((amp_t<R> *)psi_, n_tiles, prog.P, (int)rb, d_st, d_rops, d_tops, dm, two_qubit_ops, stream_);
    end_pass(e0, e1, recording_ ? &f : nullptr);
  });
  count_pass(2.0 * (double)amp_bytes() * (double)amps());
}

// Measured at 32 qubits (complex64, 68.7 GB per pass; `low` = targets at positions 0-3):
//   k = 5: VALU 15.0-16.2 ms, MFMA 13.5-14.6 ms; with low = 2 16.8-17.4 against 17.5 ms; with
//          low = 3 29.1 against 26.0 ms
//   k = 4: VALU 13.3-14.0 ms, MFMA 14.1-15.0 ms (the VALU form already streams at the in-place
//          rate); low = 3: VALU 25.4, MFMA 24.5 ms
//   k = 3: VALU 12.9-13.5, MFMA 13.4-14.7 ms; low >= 2: VALU 17.5-25.3, MFMA 16.2-19.6 ms
bool Engine::use_mfma(int k, const std::vector<uint32_t> &tpos, uint64_t n_groups) const {
  if (!c64_ || k < 3 || k > 5 || n_groups < kMfmaMinGroups || mfma_mode_ == 0) return false;
  if (mfma_mode_ == 2) return true;
  int low = 0;
  for (uint32_t p : tpos) low += p < 4;
  if (k == 5) return low != 2;
  return k == 4 ? low >= 3 : low >= 2;
}

void Engine::exec_block(const Block &b) {
  const int k = (int)b.qubits.size();
  if (k < 1 || k > 5) throw ValueError("Gates with more than 5 qubits are not supported!");
  std::vector<uint32_t> tpos(k), fixed;
  for (int l = 0; l < k; ++l) tpos[l] = pos_.at(b.qubits[l]);
  uint64_t cmask = 0;
  for (unsigned c : b.ctrl) cmask |= uint64_t(1) << pos_.at(c);
  uint64_t tmask = 0;
  for (uint32_t p : tpos) tmask |= uint64_t(1) << p;
  if (tmask & cmask) return;  // a control that is also a target is never 1 on a base index
  fixed = tpos;
  for (uint32_t p : bits_of(cmask)) fixed.push_back(p);
  BitInsert ins = make_insert(fixed);
  if (ins.count > n_) return;
  const uint64_t n_groups = amps() >> ins.count;
  const size_t D = size_t(1) << k;
  if (use_mfma(k, tpos, n_groups)) {
    const std::vector<float> table = mfma::operand_table(k, b.m.data());
    const float *d_table = (const float *)upload(table.data(), table.size() * sizeof(float));
    PassFeatures f{};
    if (recording_) f = block_features(b);
    hipEvent_t e0, e1;
    begin_pass(&e0, &e1);
    dev::apply_block_mfma((float2 *)psi_, n_groups, ins, cmask, tpos.data(), k, d_table, stream_);
    end_pass(e0, e1, recording_ ? &f : nullptr);
    count_pass(2.0 * (double)amp_bytes() * (double)(n_groups << k));
    return;
  }
  dispatch([&](auto tag) {
    using R = decltype(tag);
    std::vector<amp_t<R>> m(D * D);
    for (size_t i = 0; i < D * D; ++i) {
      m[i].x = (R)b.m[i].real();
      m[i].y = (R)b.m[i].imag();
    }
    const amp_t<R> *dm = (const amp_t<R> *)upload(m.data(), m.size() * sizeof(amp_t<R>));
    PassFeatures f{};
    if (recording_) f = block_features(b);
    hipEvent_t e0, e1;
    begin_pass(&e0, &e1);
    dev::apply_block<R>((amp_t<R> *)psi_, n_groups, ins, cmask, tpos.data(), k, dm, stream_);
    end_pass(e0, e1, recording_ ? &f : nullptr);
  });
  // algorithmic traffic: every touched amplitude is read once and written once
  count_pass(2.0 * (double)amp_bytes() * (double)(n_groups << k));
}

// ------------------------------------------------------------------------------ planner controls
void Engine::set_tile_params(unsigned tmax, unsigned rb, unsigned lmin) {
  if (tmax < 6 || tmax > (c64_ ? 13u : 12u) || (rb != 4 && rb != 5) || lmin < 1 || lmin > tmax)
    throw ValueError("tile parameters out of range");
  run();
  tile_tmax_ = tmax;
  tile_rb_ = rb;
  tile_lmin_ = lmin;
}

void Engine::set_relabel(int mode) {
  if (mode < -1 || mode > 1) throw ValueError("relabel mode must be -1 (default), 0 (off) or 1 (on)");
  relabel_mode_ = mode;
}

void Engine::set_slot_rule(int rule) {
  if (rule < 0 || rule >= kSlotRules) throw ValueError("unknown tile slot rule");
  run();
  slot_rule_ = rule;
}

void Engine::set_recording(bool on) {
  run();
  recording_ = on;
}

size_t Engine::take_records(float *features, float *ms, size_t cap) {
  run();
  sync();
  size_t k = 0;
  for (auto &r : records_) {
    if (k < cap) {
      float t = 0.f;
      QS_HIP(hipEventElapsedTime(&t, r.a, r.b));
      std::copy(r.f.f, r.f.f + kPassFeatures, features + k * kPassFeatures);
      ms[k] = t;
      ++k;
    }
    ev_free_.push_back(r.a);
    ev_free_.push_back(r.b);
  }
  records_.clear();
  return k;
}

void Engine::set_model(const std::vector<float> &w) {
  if (w.empty()) {
    model_.reset();
    return;
  }
  model_ = std::make_unique<mlp::Model>(w);
}

// ------------------------------------------------------------------------------ measurement hooks
void Engine::synchronize() { sync(); }

hipEvent_t Engine::take_event() {
  if (!ev_free_.empty()) {
    hipEvent_t e = ev_free_.back();
    ev_free_.pop_back();
    return e;
  }
  hipEvent_t e;
  QS_HIP(hipEventCreate(&e));
  return e;
}

void Engine::harvest_events() {
  sync();
  for (auto &p : ev_used_) {
    float ms = 0.f;
    QS_HIP(hipEventElapsedTime(&ms, p.first, p.second));
    stat_ms_ += ms;
    stat_timed_++;
    ev_free_.push_back(p.first);
    ev_free_.push_back(p.second);
  }
  ev_used_.clear();
}

void Engine::get_stats(uint64_t *passes, double *bytes, double *gpu_ms, uint64_t *timed) {
  harvest_events();
  *passes = stat_passes_;
  *bytes = stat_bytes_;
  *gpu_ms = stat_ms_;
  *timed = stat_timed_;
}

void Engine::reset_stats() {
  harvest_events();
  stat_passes_ = stat_timed_ = 0;
  stat_bytes_ = stat_ms_ = 0.0;
}

cplx Engine::overlap(Engine &other) {
  canonicalize();
  other.canonicalize();
  run();
  other.run();
  other.sync();
  if (other.pos_ != pos_) throw ValueError("overlap: the simulators hold different qubits or positions");
  dev::overlap(psi_, c64_, other.psi_, other.c64_, amps(), d_partials_, stream_);
  double im;
  double re = finish(dev::kReduceGrid, 2, &im);
  return cplx(re, im);
}

void Engine::compare(Engine &other, double *max_abs_diff, uint64_t *n_unequal) {
  canonicalize();
  other.canonicalize();
  run();
  other.run();
  other.sync();
  if (other.pos_ != pos_) throw ValueError("compare: the simulators hold different qubits or positions");
  dev::compare(psi_, c64_, other.psi_, other.c64_, amps(), d_partials_, stream_);
  std::vector<double> part(2 * dev::kReduceGrid);
  QS_HIP(hipMemcpyAsync(part.data(), d_partials_, part.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  sync();
  double m = 0.0, c = 0.0;
  for (int b = 0; b < dev::kReduceGrid; ++b) {
    m = std::max(m, part[2 * b]);
    c += part[2 * b + 1];
  }
  *max_abs_diff = m;
  *n_unequal = (uint64_t)c;
}

// ------------------------------------------------------------------------------ emulate_math
dev::MathArgs Engine::math_args(const std::vector<std::vector<unsigned>> &regs,
                                const std::vector<unsigned> &ctrl) const {
  if (regs.size() > (size_t)dev::kMaxRegs) throw ValueError("emulate_math: at most 8 registers");
  dev::MathArgs a{};
  a.n_regs = (int32_t)regs.size();
  for (size_t r = 0; r < regs.size(); ++r) {
    if (regs[r].size() > 32) throw ValueError("emulate_math: registers hold at most 32 qubits");
    a.width[r] = (uint8_t)regs[r].size();
    for (size_t b = 0; b < regs[r].size(); ++b) a.pos[r][b] = (uint8_t)pos_.at(regs[r][b]);
  }
  a.cmask = mask_of(ctrl);
  return a;
}

void Engine::math_apply(dev::MathArgs args, const int32_t *d_table, bool known_bijective) {
  subspace_unknown();  // amplitudes may move anywhere
  const uint64_t N = amps();
  QS_HIP(hipMemsetAsync(d_flags_, 0, sizeof(int) * 2, stream_));
  bool injective = known_bijective;
  if (!injective) {
    const size_t words = (N + 31) / 32;
    uint32_t *bitmap = nullptr;
    QS_HIP(hipMalloc(&bitmap, words * sizeof(uint32_t)));
    QS_HIP(hipMemsetAsync(bitmap, 0, words * sizeof(uint32_t), stream_));
    dev::math_injective(N, args, d_table, bitmap, d_flags_ + 1, d_flags_, stream_);
    QS_HIP(hipMemcpyAsync(h_flags_, d_flags_, sizeof(int) * 2, hipMemcpyDeviceToHost, stream_));
    sync();
    QS_HIP(hipFree(bitmap));
    if (h_flags_[0]) throw RuntimeError("emulate_math: a register value does not fit a C++ int");
    injective = h_flags_[1] == 0;
  }
  void *dst = scratch(0);
  dispatch([&](auto tag) {
    using R = decltype(tag);
    if (injective) {
      dev::math_permute<R>((const amp_t<R> *)psi_, (amp_t<R> *)dst, N, args, d_table, d_flags_, stream_);
    } else {
      QS_HIP(hipMemsetAsync(dst, 0, N * amp_bytes(), stream_));
      dev::math_accumulate<R>((const amp_t<R> *)psi_, (amp_t<R> *)dst, N, args, d_table, d_flags_, stream_);
    }
  });
  QS_HIP(hipMemcpyAsync(h_flags_, d_flags_, sizeof(int), hipMemcpyDeviceToHost, stream_));
  sync();
  if (h_flags_[0]) throw RuntimeError("emulate_math: a register value does not fit a C++ int");
  std::swap(psi_, buf_[0]);
  norm_changed();
  subspace_unknown();
}

// simulator.hpp:224-290 for the opcodes of include/projectq_amd.h
void Engine::emulate_math_op(int opcode, const std::vector<int64_t> &params,
                             const std::vector<std::vector<unsigned>> &regs,
                             const std::vector<unsigned> &ctrl) {
  canonicalize();  // colliding amplitudes add in the reference's index order
  for (auto &r : regs) check_known(r, "emulate_math");
  check_known(ctrl, "emulate_math");
  dev::MathArgs a = math_args(regs, ctrl);
  a.opcode = opcode;
  a.a = params.size() > 0 ? params[0] : 0;
  a.N = params.size() > 1 ? params[1] : 1;
  // x -> x + a wraps modulo 2^width on every register independently: a bijection
  math_apply(a, nullptr, opcode == 0);
}

// simulator.hpp:224-269 with a host function: evaluated once per distinct tuple of register
// values into a lookup table, then applied as a device permutation.
void Engine::emulate_math_callback(const std::function<bool(int32_t *, size_t)> &fn,
                                   const std::vector<std::vector<unsigned>> &regs,
                                   const std::vector<unsigned> &ctrl) {
  canonicalize();  // colliding amplitudes add in the reference's index order
  for (auto &r : regs) check_known(r, "emulate_math");
  check_known(ctrl, "emulate_math");
  dev::MathArgs a = math_args(regs, ctrl);
  a.opcode = -1;
  unsigned bits = 0;
  for (auto &r : regs) bits += (unsigned)r.size();
  if (bits > 30) throw ValueError("emulate_math: a Python function over more than 30 register qubits is not supported");
  const size_t domain = size_t(1) << bits, nr = regs.size();
  std::vector<int32_t> table(domain * std::max<size_t>(nr, 1));
  std::vector<int32_t> v(std::max<size_t>(nr, 1));
  for (size_t x = 0; x < domain; ++x) {
    unsigned shift = 0;
    for (size_t r = 0; r < nr; ++r) {
      uint64_t w = regs[r].size();
      v[r] = (int32_t)(uint32_t)((x >> shift) & ((uint64_t(1) << w) - 1));
      shift += (unsigned)w;
    }
    if (!fn(v.data(), nr)) throw CallbackError("emulate_math: the Python function raised");
    std::copy(v.begin(), v.begin() + nr, table.begin() + x * nr);
  }
  int32_t *d_table = nullptr;
  QS_HIP(hipMalloc(&d_table, table.size() * sizeof(int32_t)));
  try {
    QS_HIP(hipMemcpyAsync(d_table, table.data(), table.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream_));
    math_apply(a, d_table, false);
  } catch (...) {
    (void)hipStreamSynchronize(stream_);
    (void)hipFree(d_table);
    throw;
  }
  QS_HIP(hipFree(d_table));
}

// ------------------------------------------------------------------------------ Pauli strings
// A Pauli string applied left to right (simulator.hpp:538-550) equals omega * X^x Z^z:
// X_q: x ^= q;  Z_q: omega *= (-1)^{x_q}, z ^= q;  Y_q = i X_q Z_q: omega *= i (-1)^{x_q},
// x ^= q, z ^= q.  (P psi)_i = omega (-1)^{popcount((i^x)&z)} psi_{i^x}.
void Engine::pauli_encode(const PauliString &t, const std::vector<unsigned> &ids, uint64_t *x,
                          uint64_t *z, cplx *omega) const {
  std::vector<std::pair<uint32_t, char>> ops;
  for (auto &op : t.ops) {
    if (op.first >= ids.size()) throw ValueError("qubit_operator acts on more qubits than contained in the qureg.");
    if (op.second != 'X' && op.second != 'Y' && op.second != 'Z') throw ValueError("Pauli operator must be X, Y or Z");
    ops.push_back({pos_.at(ids[op.first]), op.second});
  }
  encode_pauli(ops, x, z, omega);
}

// simulator.hpp:292-322: sum_t c_t Re<psi|P_t|psi>.  A state that fits in the last-level cache
// is read once for all strings (one FP64 accumulator per thread: the coefficients and phases
// are folded into each string); a larger state is read once per string so every pass streams.
double Engine::get_expectation_value(const std::vector<PauliString> &terms,
                                     const std::vector<double> &coeffs,
                                     const std::vector<unsigned> &ids) {
  run();
  check_known(ids, "get_expectation_value");
  std::vector<dev::PauliTerm> pt(terms.size());
  std::vector<cplx> omega(terms.size());
  for (size_t t = 0; t < terms.size(); ++t) pauli_encode(terms[t], ids, &pt[t].x, &pt[t].z, &omega[t]);
  const bool cached = amps() * amp_bytes() <= (uint64_t(64) << 20);
  if (cached && !pt.empty()) {
    for (size_t t = 0; t < pt.size(); ++t) {
      if (!dev::pauli_hermitian(pt[t].x, pt[t].z, omega[t])) {
        pt[t] = dev::PauliTerm{0, 0, 0.0, 0.0};  // Re<psi|P|psi> = 0 for an anti-Hermitian string
        continue;
      }
      const cplx c = coeffs[t] * omega[t];
      pt[t].c_re = c.real();
      pt[t].c_im = c.imag();
    }
    const dev::PauliTerm *d_terms = (const dev::PauliTerm *)upload(pt.data(), pt.size() * sizeof(dev::PauliTerm));
    dispatch([&](auto tag) {
      using R = decltype(tag);
      dev::pauli_energy<R>((const amp_t<R> *)psi_, amps(), d_terms, (int)pt.size(), d_partials_, stream_);
    });
    return finish(dev::kReduceGrid, 1);
  }
  double total = 0.0;
  for (size_t t = 0; t < terms.size(); ++t) {
    if (!dev::pauli_hermitian(pt[t].x, pt[t].z, omega[t])) continue;  // Re<psi|P|psi> = 0
    double delta;
    if (pt[t].x == 0 && pt[t].z == 0) {
      delta = omega[t].real() * masked_norm(0, 0);
    } else {
      dispatch([&](auto tag) {
        using R = decltype(tag);
        dev::pauli_expect<R>((const amp_t<R> *)psi_, amps(), pt[t].x, pt[t].z, d_partials_, stream_);
      });
      double si;
      double sr = finish(dev::kReduceGrid, 2, &si);
      delta = (omega[t] * cplx(sr, si)).real();
    }
    total += coeffs[t] * delta;
  }
  return total;
}

// simulator.hpp:324-350: psi <- sum_t c_t P_t psi in one out-of-place pass.
void Engine::apply_qubit_operator(const std::vector<PauliString> &terms, const std::vector<cplx> &coeffs,
                                  const std::vector<unsigned> &ids) {
  subspace_unknown();  // amplitudes may move anywhere
  run();
  check_known(ids, "apply_qubit_operator");
  std::vector<dev::PauliTerm> pt(terms.size());
  for (size_t t = 0; t < terms.size(); ++t) {
    cplx w;
    pauli_encode(terms[t], ids, &pt[t].x, &pt[t].z, &w);
    cplx c = coeffs[t] * w;
    pt[t].c_re = c.real();
    pt[t].c_im = c.imag();
  }
  const dev::PauliTerm *d_terms =
      pt.empty() ? nullptr : (const dev::PauliTerm *)upload(pt.data(), pt.size() * sizeof(dev::PauliTerm));
  void *dst = scratch(0);
  dispatch([&](auto tag) {
    using R = decltype(tag);
    dev::pauli_sum<R>((const amp_t<R> *)psi_, (amp_t<R> *)dst, nullptr, amps(), d_terms, (int)pt.size(), 1.0,
                      0.0, false, 0, d_partials_, stream_);
  });
  std::swap(psi_, buf_[0]);
  norm_changed();
  subspace_unknown();
}

// simulator.hpp:386-438: scaled, truncated Taylor series of exp(-i H t) on the controlled
// subspace; each Taylor term is one pass that also accumulates the output and the update norm.
void Engine::emulate_time_evolution(const std::vector<PauliString> &terms, const std::vector<double> &coeffs,
                                    double time, const std::vector<unsigned> &ids,
                                    const std::vector<unsigned> &ctrl) {
  subspace_unknown();  // amplitudes may move anywhere
  run();
  check_known(ids, "emulate_time_evolution");
  check_known(ctrl, "emulate_time_evolution");
  const cplx I(0.0, 1.0);
  double tr = 0.0, op_nrm = 0.0;
  std::vector<dev::PauliTerm> pt;
  for (size_t t = 0; t < terms.size(); ++t) {
    if (terms[t].ops.empty()) {
      tr += coeffs[t];
      continue;
    }
    dev::PauliTerm p;
    cplx w;
    pauli_encode(terms[t], ids, &p.x, &p.z, &w);
    cplx c = coeffs[t] * w;
    p.c_re = c.real();
    p.c_im = c.imag();
    pt.push_back(p);
    op_nrm += std::abs(coeffs[t]);
  }
  const unsigned steps = (unsigned)(std::abs(time) * op_nrm + 1.);
  const cplx correction = std::exp(-time * I * tr / (double)steps);
  const uint64_t cmask = mask_of(ctrl), N = amps();
  const dev::PauliTerm *d_terms =
      pt.empty() ? nullptr : (const dev::PauliTerm *)upload(pt.data(), pt.size() * sizeof(dev::PauliTerm));
  void *out = scratch(0);
  scratch(1);  // buf_[1] ping-pongs with psi_ as the Taylor term; psi_, buf_[0], buf_[1] stay distinct
  QS_HIP(hipMemcpyAsync(out, psi_, N * amp_bytes(), hipMemcpyDeviceToDevice, stream_));
  for (unsigned s = 0; s < steps; ++s) {
    double change = 1.0;
    for (unsigned k = 0; change > 1.e-12; ++k) {
      const cplx f = (-time * I) / double(steps * (k + 1));
      dispatch([&](auto tag) {
        using R = decltype(tag);
        dev::pauli_sum<R>((const amp_t<R> *)psi_, (amp_t<R> *)buf_[1], (amp_t<R> *)out, N, d_terms,
                          (int)pt.size(), f.real(), f.imag(), true, cmask, d_partials_, stream_);
      });
      change = std::sqrt(finish(dev::kReduceGrid, 1));
      std::swap(psi_, buf_[1]);
    }
    dispatch([&](auto tag) {
      using R = decltype(tag);
      dev::scale_masked<R>((amp_t<R> *)out, N, cmask, correction.real(), correction.imag(), stream_);
    });
    QS_HIP(hipMemcpyAsync(psi_, out, N * amp_bytes(), hipMemcpyDeviceToDevice, stream_));
  }
  sync();
  norm_changed();
  subspace_unknown();
}

// ------------------------------------------------------------------------------ queries
// simulator.hpp:352-368
double Engine::get_probability(const std::vector<uint8_t> &bits, const std::vector<unsigned> &ids) {
  run();
  for (unsigned q : ids)
    if (!pos_.count(q))
      throw RuntimeError("get_probability(): Unknown qubit id. Please make sure you have called eng.flush().");
  uint64_t mask = 0, val = 0;
  for (size_t i = 0; i < ids.size(); ++i) {
    mask |= uint64_t(1) << pos_[ids[i]];
    val |= uint64_t(bits[i] ? 1 : 0) << pos_[ids[i]];
  }
  return masked_norm(mask, val & mask);
}

// simulator.hpp:370-384
cplx Engine::get_amplitude(const std::vector<uint8_t> &bits, const std::vector<unsigned> &ids) {
  run();
  uint64_t chk = 0, index = 0;
  for (size_t i = 0; i < ids.size(); ++i) {
    auto it = pos_.find(ids[i]);
    if (it == pos_.end()) break;
    chk |= uint64_t(1) << it->second;
    index |= uint64_t(bits[i] ? 1 : 0) << it->second;
  }
  if (chk + 1 != amps())
    throw RuntimeError(
        "The second argument to get_amplitude() must be a permutation of all allocated qubits. Please make "
        "sure you have called eng.flush().");
  if (c64_) {
    float a[2];
    QS_HIP(hipMemcpyAsync(a, (char *)psi_ + index * 8, 8, hipMemcpyDeviceToHost, stream_));
    sync();
    return cplx(a[0], a[1]);
  }
  double a[2];
  QS_HIP(hipMemcpyAsync(a, (char *)psi_ + index * 16, 16, hipMemcpyDeviceToHost, stream_));
  sync();
  return cplx(a[0], a[1]);
}

// simulator.hpp:440-454
void Engine::set_wavefunction(const double *wf, size_t n_amps, const std::vector<unsigned> &ordering) {
  subspace_unknown();  // amplitudes may move anywhere
  run();
  bool ok = pos_.size() == ordering.size();
  for (unsigned q : ordering) ok = ok && pos_.count(q);
  if (!ok)
    throw RuntimeError(
        "set_wavefunction(): Invalid mapping provided. Please make sure all qubits have been allocated "
        "previously (call eng.flush()).");
  if (n_amps != (uint64_t(1) << ordering.size()))
    throw ValueError("set_wavefunction(): the wavefunction must hold 2^len(ordering) amplitudes");
  for (size_t i = 0; i < ordering.size(); ++i) pos_[ordering[i]] = (unsigned)i;
  ref_pos_ = pos_;
  if (c64_) {
    std::vector<float> f(2 * n_amps);
    for (size_t i = 0; i < 2 * n_amps; ++i) f[i] = (float)wf[i];
    QS_HIP(hipMemcpyAsync(psi_, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, stream_));
    sync();
  } else {
    QS_HIP(hipMemcpyAsync(psi_, wf, n_amps * 16, hipMemcpyHostToDevice, stream_));
    sync();
  }
  norm_changed();
  subspace_unknown();
}

// simulator.hpp:456-485
void Engine::collapse_wavefunction(const std::vector<unsigned> &ids, const std::vector<uint8_t> &values) {
  run();
  if (ids.size() != values.size()) throw ValueError("collapse_wavefunction(): ids and values size mismatch");
  for (unsigned q : ids)
    if (!pos_.count(q))
      throw RuntimeError(
          "collapse_wavefunction(): Unknown qubit id(s) provided. Try calling eng.flush() before invoking "
          "this function.");
  uint64_t mask = 0, val = 0;
  for (size_t i = 0; i < ids.size(); ++i) {
    mask |= uint64_t(1) << pos_[ids[i]];
    val |= uint64_t(values[i] ? 1 : 0) << pos_[ids[i]];
  }
  val &= mask;
  const double p = ((val ^ zval_) & mask & zmask_) ? 0.0 : masked_norm(mask | zmask_, val | (zval_ & ~mask));
  if (p < 1.e-12) throw RuntimeError("collapse_wavefunction(): Invalid collapse! Probability is ~0.");
  const double scale = 1. / std::sqrt(p);
  const BitInsert sub = make_insert(bits_of(zmask_));
  dispatch([&](auto tag) {
    using R = decltype(tag);
    dev::collapse<R>((amp_t<R> *)psi_, amps() >> sub.count, sub, zval_, mask, val, scale, stream_);
  });
  subspace_fix(mask, val);
  norm_changed();
}

// simulator.hpp:529-532
void Engine::cheat_sizes(size_t *n_qubits, size_t *n_amps) {
  canonicalize();
  sync();
  *n_qubits = pos_.size();
  *n_amps = amps();
}

void Engine::cheat(uint32_t *ids, uint32_t *positions, double *state) {
  canonicalize();
  size_t k = 0;
  for (auto &kv : pos_) {
    ids[k] = kv.first;
    positions[k] = kv.second;
    ++k;
  }
  const uint64_t N = amps();
  if (c64_) {
    std::vector<float> f(2 * N);
    QS_HIP(hipMemcpyAsync(f.data(), psi_, N * 8, hipMemcpyDeviceToHost, stream_));
    sync();
    for (size_t i = 0; i < 2 * N; ++i) state[i] = f[i];
  } else {
    QS_HIP(hipMemcpyAsync(state, psi_, N * 16, hipMemcpyDeviceToHost, stream_));
    sync();
  }
}


// The state stays where it is: the pointer is valid until the next call on this simulator that
// applies gates, allocates or frees qubits (those may move the state to the other buffer).
void Engine::device_state(void **ptr, uint64_t *n_amps, int *precision) {
  canonicalize();
  sync();
  *ptr = psi_;
  *n_amps = amps();
  *precision = c64_ ? 1 : 0;
  norm_changed();  // the caller may write through the pointer
  subspace_unknown();
  pointer_out_ = true;
}

void Engine::qubit_map(uint32_t *ids, uint32_t *positions, size_t capacity, size_t *n_qubits) {
  canonicalize();
  size_t k = 0;
  for (auto &kv : pos_) {
    if (k < capacity) {
      ids[k] = kv.first;
      positions[k] = kv.second;
    }
    ++k;
  }
  *n_qubits = k;
}

namespace {
// Checkpoint file: header, the qubit map, then the amplitudes (interleaved re/im, little-endian)
// in the simulator's precision.
constexpr char kStateMagic[8] = {'Q', 'S', 'A', 'M', 'D', 'S', 'T', '1'};
struct StateHeader {
  char magic[8];
  uint32_t version;    // 1
  uint32_t precision;  // 0 complex128, 1 complex64
  uint32_t n_qubits;
  uint32_t reserved;
  uint64_t n_amps;
};
constexpr size_t kStageBytes = size_t(64) << 20;  // pinned staging buffer (two of them)

struct PinnedPair {
  void *p[2] = {nullptr, nullptr};
  PinnedPair() {
    for (auto &b : p) QS_HIP(hipHostMalloc(&b, kStageBytes, hipHostMallocDefault));
  }
  ~PinnedPair() {
    for (auto &b : p)
      if (b) (void)hipHostFree(b);
  }
};
struct File {
  FILE *f;
  File(const char *path, const char *mode) : f(std::fopen(path, mode)) {
    if (!f) throw RuntimeError(std::string("cannot open state file ") + path);
  }
  ~File() {
    if (f) std::fclose(f);
  }
};
}  // namespace

void Engine::save_state(const char *path) {
  canonicalize();
  sync();
  File file(path, "wb");
  StateHeader h{};
  std::memcpy(h.magic, kStateMagic, 8);
  h.version = 1;
  h.precision = c64_ ? 1 : 0;
  h.n_qubits = n_;
  h.n_amps = amps();
  if (std::fwrite(&h, sizeof h, 1, file.f) != 1) throw RuntimeError("save_state: write failed");
  for (auto &kv : pos_) {
    const uint32_t rec[2] = {kv.first, kv.second};
    if (std::fwrite(rec, sizeof rec, 1, file.f) != 1) throw RuntimeError("save_state: write failed");
  }
  // device -> pinned (async, alternating buffers) -> file
  PinnedPair stage;
  const uint64_t total = amps() * amp_bytes();
  const char *src = static_cast<const char *>(psi_);
  hipEvent_t done[2];
  for (auto &e : done) QS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  uint64_t off = 0;
  int k = 0;
  uint64_t pending = 0;  // bytes in flight into stage.p[k ^ 1]
  try {
    while (off < total || pending) {
      uint64_t len = 0;
      if (off < total) {
        len = std::min<uint64_t>(kStageBytes, total - off);
        QS_HIP(hipMemcpyAsync(stage.p[k], src + off, len, hipMemcpyDeviceToHost, stream_));
        QS_HIP(hipEventRecord(done[k], stream_));
        off += len;
      }
      if (pending) {
        QS_HIP(hipEventSynchronize(done[k ^ 1]));
        if (std::fwrite(stage.p[k ^ 1], 1, pending, file.f) != pending) throw RuntimeError("save_state: write failed");
      }
      pending = len;
      k ^= 1;
    }
  } catch (...) {
    (void)hipStreamSynchronize(stream_);
    for (auto &e : done) (void)hipEventDestroy(e);
    throw;
  }
  for (auto &e : done) QS_HIP(hipEventDestroy(e));
}

void Engine::load_state(const char *path) {
  run();
  File file(path, "rb");
  StateHeader h{};
  if (std::fread(&h, sizeof h, 1, file.f) != 1 || std::memcmp(h.magic, kStateMagic, 8) != 0 || h.version != 1)
    throw ValueError("load_state: not a state file of this simulator");
  if (h.n_qubits > 40 || h.n_amps != (uint64_t(1) << h.n_qubits) || h.precision > 1)
    throw ValueError("load_state: corrupt header");
  // the whole file must be there before the qubit map or the state is touched
  {
    const long here = std::ftell(file.f);
    if (here < 0 || std::fseek(file.f, 0, SEEK_END) != 0) throw RuntimeError("load_state: cannot size the file");
    const long size = std::ftell(file.f);
    if (size < 0 || std::fseek(file.f, here, SEEK_SET) != 0) throw RuntimeError("load_state: cannot size the file");
    const uint64_t want = sizeof(StateHeader) + uint64_t(8) * h.n_qubits + h.n_amps * (h.precision == 1 ? 8u : 16u);
    if ((uint64_t)size != want) throw ValueError("load_state: the file is truncated or has trailing bytes");
  }
  std::map<unsigned, unsigned> map;
  std::vector<bool> seen(h.n_qubits, false);
  for (uint32_t i = 0; i < h.n_qubits; ++i) {
    uint32_t rec[2];
    if (std::fread(rec, sizeof rec, 1, file.f) != 1) throw ValueError("load_state: truncated qubit map");
    if (rec[1] >= h.n_qubits || seen[rec[1]] || map.count(rec[0])) throw ValueError("load_state: corrupt qubit map");
    seen[rec[1]] = true;
    map[rec[0]] = rec[1];
  }
  if (!pos_.empty()) {
    bool same = pos_.size() == map.size();
    for (auto &kv : map) same = same && pos_.count(kv.first);
    if (!same) throw RuntimeError("load_state: the simulator holds other qubits than the file");
  } else {
    std::vector<unsigned> by_pos(h.n_qubits);
    for (auto &kv : map) by_pos[kv.second] = kv.first;
    for (unsigned id : by_pos) allocate_qubit(id);
  }
  pos_ = map;
  ref_pos_ = map;
  const bool file_c64 = h.precision == 1;
  const size_t file_amp = file_c64 ? 8 : 16;
  const uint64_t total = h.n_amps * file_amp;
  PinnedPair stage;
  char *dst = static_cast<char *>(psi_);
  uint64_t off = 0;
  int k = 0;
  sync();
  while (off < total) {
    const uint64_t len = std::min<uint64_t>(kStageBytes - kStageBytes % 16, total - off);
    // the previous copy out of stage.p[k] must be done before it is refilled
    QS_HIP(hipStreamSynchronize(stream_));
    if (std::fread(stage.p[k], 1, len, file.f) != len) throw ValueError("load_state: truncated amplitudes");
    const uint64_t first_amp = off / file_amp, n = len / file_amp;
    if (file_c64 == c64_) {
      QS_HIP(hipMemcpyAsync(dst + off, stage.p[k], len, hipMemcpyHostToDevice, stream_));
    } else {
      // convert precision in place on the host side of the other buffer, then copy
      void *other = stage.p[k ^ 1];
      if (c64_) {
        const double *in = static_cast<const double *>(stage.p[k]);
        float *out = static_cast<float *>(other);
        for (uint64_t i = 0; i < 2 * n; ++i) out[i] = (float)in[i];
        QS_HIP(hipMemcpyAsync(dst + first_amp * 8, other, n * 8, hipMemcpyHostToDevice, stream_));
      } else {
        // complex64 file into a complex128 simulator: two halves so the output fits the buffer
        const float *in = static_cast<const float *>(stage.p[k]);
        double *out = static_cast<double *>(other);
        const uint64_t half = n / 2;
        for (uint64_t part = 0; part < 2; ++part) {
          const uint64_t a = part ? half : 0, b = part ? n : half;
          for (uint64_t i = 2 * a; i < 2 * b; ++i) out[i - 2 * a] = (double)in[i];
          QS_HIP(hipMemcpyAsync(dst + (first_amp + a) * 16, other, (b - a) * 16, hipMemcpyHostToDevice, stream_));
          QS_HIP(hipStreamSynchronize(stream_));
        }
      }
    }
    off += len;
    k ^= 1;
  }
  sync();
  norm_changed();
  subspace_unknown();
}

void Engine::set_wavefunction_device(const void *src, uint64_t n_amps, int src_precision,
                                     const std::vector<unsigned> &ordering) {
  subspace_unknown();  // amplitudes may move anywhere
  run();
  bool ok = pos_.size() == ordering.size();
  for (unsigned q : ordering) ok = ok && pos_.count(q);
  if (!ok)
    throw RuntimeError(
        "set_wavefunction(): Invalid mapping provided. Please make sure all qubits have been allocated "
        "previously (call eng.flush()).");
  if (n_amps != (uint64_t(1) << ordering.size()))
    throw ValueError("set_wavefunction(): the wavefunction must hold 2^len(ordering) amplitudes");
  if (src_precision != 0 && src_precision != 1) throw ValueError("set_wavefunction_device: unknown precision");
  for (size_t i = 0; i < ordering.size(); ++i) pos_[ordering[i]] = (unsigned)i;
  ref_pos_ = pos_;
  const bool src_c64 = src_precision == 1;
  if (src_c64 == c64_) {
    QS_HIP(hipMemcpyAsync(psi_, src, n_amps * amp_bytes(), hipMemcpyDeviceToDevice, stream_));
  } else if (c64_) {
    dev::convert_amps<double, float>(static_cast<const amp_t<double> *>(src), static_cast<amp_t<float> *>(psi_),
                                     n_amps, stream_);
  } else {
    dev::convert_amps<float, double>(static_cast<const amp_t<float> *>(src), static_cast<amp_t<double> *>(psi_),
                                     n_amps, stream_);
  }
  sync();
  norm_changed();
  subspace_unknown();
}

}  // namespace qs
