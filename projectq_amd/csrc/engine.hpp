// The simulator engine: owns the resident HBM state, the qubit map, the RNG and the gate queue,
// and implements every method of the reference's class Simulator
// (projectq/backends/_sim/_cppkernels/simulator.hpp:37-579) on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <random>
#include <utility>
#include <vector>

#include "common.hpp"
#include "fusion.hpp"
#include "kernels.hpp"
#include "mlp.hpp"
#include "planner.hpp"

namespace qs {

struct PauliString {
  std::vector<std::pair<uint32_t, char>> ops;  // (local index into ids, 'X'|'Y'|'Z')
};

class Engine {
 public:
  Engine(uint32_t seed, int precision, int fusion_mode, int planner, unsigned capacity_hint);
  ~Engine();
  Engine(const Engine &) = delete;
  Engine &operator=(const Engine &) = delete;

  void allocate_qubit(unsigned id);
  void deallocate_qubit(unsigned id);
  bool get_classical_value(unsigned id, double tol);
  bool is_classical(unsigned id, double tol);
  std::vector<uint8_t> measure_qubits(const std::vector<unsigned> &ids);
  void apply_controlled_gate(Gate g);
  void emulate_math_op(int opcode, const std::vector<int64_t> &params,
                       const std::vector<std::vector<unsigned>> &regs,
                       const std::vector<unsigned> &ctrl);
  void emulate_math_callback(const std::function<bool(int32_t *, size_t)> &fn,
                             const std::vector<std::vector<unsigned>> &regs,
                             const std::vector<unsigned> &ctrl);
  double get_expectation_value(const std::vector<PauliString> &terms,
                               const std::vector<double> &coeffs,
                               const std::vector<unsigned> &ids);
  void apply_qubit_operator(const std::vector<PauliString> &terms, const std::vector<cplx> &coeffs,
                            const std::vector<unsigned> &ids);
  void emulate_time_evolution(const std::vector<PauliString> &terms,
                              const std::vector<double> &coeffs, double time,
                              const std::vector<unsigned> &ids, const std::vector<unsigned> &ctrl);
  double get_probability(const std::vector<uint8_t> &bits, const std::vector<unsigned> &ids);
  cplx get_amplitude(const std::vector<uint8_t> &bits, const std::vector<unsigned> &ids);
  void set_wavefunction(const double *wf, size_t n_amps, const std::vector<unsigned> &ordering);
  void collapse_wavefunction(const std::vector<unsigned> &ids, const std::vector<uint8_t> &values);
  void run();
  void cheat_sizes(size_t *n_qubits, size_t *n_amps);
  void cheat(uint32_t *ids, uint32_t *positions, double *state);
  // Extensions: the state in device memory without a host copy, checkpoint/restore to a file,
  // and set_wavefunction from device memory.
  void device_state(void **ptr, uint64_t *n_amps, int *precision);
  void qubit_map(uint32_t *ids, uint32_t *positions, size_t capacity, size_t *n_qubits);
  void save_state(const char *path);
  void load_state(const char *path);
  void set_wavefunction_device(const void *src, uint64_t n_amps, int src_precision,
                               const std::vector<unsigned> &ordering);

  // execute one fused block (also used by the planners)
  void exec_block(const Block &b);

  // planner controls (C-ABI extensions)
  void set_tile_params(unsigned tmax, unsigned rb, unsigned lmin);
  void set_slot_rule(int rule);
  void set_relabel(int mode);
  void set_recording(bool on);
  size_t take_records(float *features, float *ms, size_t cap);
  void set_model(const std::vector<float> &w);
  struct PlanInfo {
    int kind = -1;  // 0 greedy blocks, 1 tile passes
    unsigned tmax = 0, rb = 0, lmin = 0, passes = 0, candidates = 0;
    int rule = 0;   // tile slot rule (SlotRule)
    double predicted_ms = 0.0;
    unsigned relabels = 0;  // tile passes that relabel qubits at their stores
  };
  const PlanInfo &last_plan() const { return last_plan_; }

  // measurement hooks (C-ABI extensions)
  void synchronize();
  void set_profiling(bool on) { profiling_ = on; }
  void get_stats(uint64_t *passes, double *bytes, double *gpu_ms, uint64_t *timed);
  void reset_stats();
  cplx overlap(Engine &other);
  // max |a_i - b_i| and the number of amplitudes whose stored values differ (same qubits at the
  // same positions required)
  void compare(Engine &other, double *max_abs_diff, uint64_t *n_unequal);
  unsigned num_qubits() const { return n_; }
  unsigned position(unsigned id) const { return pos_.at(id); }
  hipStream_t stream() const { return stream_; }
  bool is_c64() const { return c64_; }

 private:
  template <class F>
  void dispatch(F &&f) {
    if (c64_) f(float{});
    else f(double{});
  }
  size_t amp_bytes() const { return c64_ ? 8 : 16; }
  uint64_t amps() const { return uint64_t(1) << n_; }
  void ensure_capacity(uint64_t need_amps);
  void *scratch(int which);  // full-capacity scratch buffer 0 or 1 (lazy)
  void release_scratch();
  const void *upload(const void *host, size_t bytes);
  double finish(int n_partials, int width, double *second = nullptr);
  void sync();
  void check_known(const std::vector<unsigned> &ids, const char *what) const;
  uint64_t mask_of(const std::vector<unsigned> &ids) const;
  void pauli_encode(const PauliString &t, const std::vector<unsigned> &ids, uint64_t *x,
                    uint64_t *z, cplx *omega) const;
  dev::MathArgs math_args(const std::vector<std::vector<unsigned>> &regs,
                          const std::vector<unsigned> &ctrl) const;
  void math_apply(dev::MathArgs args, const int32_t *d_table, bool known_bijective);
  double masked_norm(uint64_t mask, uint64_t val);
  void flush_window();
  // complex64 norm guard: FP32-rounded gate matrices are unitary only to ~1e-8 (the FP32 H scales
  // the norm by 1 - 3.4e-8), so over thousands of gates the state's norm drifts from its exact
  // value.  While only unitary gates run, the FP64 norm taken before them is the exact target.
  // Each gate adds a bound on its norm change, ||M32^H M32 - I||_F of its FP32-rounded matrix; when
  // the bound accumulated since the last check reaches kNormDriftBudget, the state is rescaled to
  // the target after the execution (one norm pass and one scaling pass).
  static constexpr double kNormDriftBudget = 2.5e-7;
  double norm_ref_ = 1.0;      // norm^2 of the state before the unitary gates since the last check
  bool norm_known_ = true;     // norm_ref_ is valid (the empty state |> has norm 1)
  double drift_bound_ = 0.0;   // bound on the relative norm change of the gates since the check
  bool queued_nonunitary_ = false;  // a gate queued since the last execution is not unitary
  double queued_drift_ = 0.0;       // drift bound of the gates queued since the last execution
  void norm_changed() {
    norm_known_ = false;
    drift_bound_ = 0.0;
  }
  void before_unitary_passes();
  void after_passes();
  // Known-zero amplitudes: every amplitude i with (i & zmask_) != zval_ is exactly zero (qubits
  // fixed by a measurement or a collapse, or freshly allocated and not touched by a non-diagonal
  // gate since).  Measurements then read and write only that subspace: after All(Measure) the k-th
  // qubit's measurement touches 2^(n-k+1) amplitudes.  A non-diagonal gate frees its targets; an
  // operation that can move amplitudes anywhere, or a handed-out device pointer, frees all bits.
  uint64_t zmask_ = 0, zval_ = 0;
  bool pointer_out_ = false;  // device_state handed the state out; it may be written until the
                              // next call that applies gates, allocates or frees qubits
  void subspace_unknown() { zmask_ = zval_ = 0; }
  void subspace_fix(uint64_t mask, uint64_t val) {
    if (pointer_out_) return;
    zmask_ |= mask;
    zval_ = (zval_ & ~mask) | (val & mask);
  }
  void plan_window(const std::vector<Gate> &gates, std::vector<Block> *blocks, std::vector<TileProgram> *progs,
                   unsigned *rb, PlanInfo *info);
  unsigned tile_tmax_ = 13, tile_rb_ = 5;  // tile shape limits (PROJECTQ_AMD_TILE="t:rb" overrides)
  unsigned tile_bits(unsigned tmax) const;
  std::vector<Block> greedy_blocks(const std::vector<Gate> &gates) const;
  std::vector<PGate> resolve(const std::vector<Gate> &gates) const;
  std::vector<TileProgram> plan_tiles(const std::vector<PGate> &pg, unsigned tmax, unsigned rb, unsigned lmin,
                                      int rule, bool relabel, unsigned *t_used) const;
  int slot_rule_ = kSlotLookahead;  // rule of QS_PLANNER_TILE
  // qubit relabelling at tile stores (schedule_tiles): -1 the default (on for the learned
  // planner, off for QS_PLANNER_TILE), 0 off, 1 on; PROJECTQ_AMD_RELABEL=0|1 sets it
  int relabel_mode_ = -1;
  // window planners: CNOTs between 1-qubit gates on their target become H, CZ, H (cnot_to_cz) in
  // windows without 2-qubit gates; PROJECTQ_AMD_CZ=0 keeps them as X ops
  bool cnot_cz_ = true;
  void apply_relabel(const std::vector<std::pair<uint32_t, uint32_t>> &swaps);
  void exec_tile_program(const TileProgram &prog, unsigned rb);
  PassFeatures block_features(const Block &b) const;
  void begin_pass(hipEvent_t *e0, hipEvent_t *e1);
  void end_pass(hipEvent_t e0, hipEvent_t e1, const PassFeatures *f);
  unsigned tile_lmin_ = 5;
  // complex64 dense blocks on the matrix cores: 0 never (PROJECTQ_AMD_MFMA=0), 1 where measured
  // faster than the VALU form (the default; use_mfma), 2 every block of 3-5 qubits
  // (PROJECTQ_AMD_MFMA=all)
  int mfma_mode_ = 1;
  static constexpr uint64_t kMfmaMinGroups = 4096;
  // learned planner: candidates with more than this many times the fewest passes are not scored
  static constexpr size_t kMaxPassRatio = 3;
  bool use_mfma(int k, const std::vector<uint32_t> &tpos, uint64_t n_groups) const;
  bool recording_ = false;
  struct PassRecord {
    PassFeatures f;
    hipEvent_t a, b;
  };
  std::vector<PassRecord> records_;
  std::unique_ptr<mlp::Model> model_;
  PlanInfo last_plan_;

  bool c64_;
  int fusion_mode_, planner_kind_;
  int device_ = 0;
  unsigned n_ = 0;
  unsigned cap_hint_ = 0;
  uint64_t cap_ = 0;  // amplitudes the buffers hold
  void *psi_ = nullptr;
  void *buf_[2] = {nullptr, nullptr};
  std::map<unsigned, unsigned> pos_;
  // The reference's qubit layout (simulator.hpp: a new qubit takes the next position, a freed one
  // closes the gap).  Relabelling tile stores move qubits away from it (pos_ is the layout of the
  // state in HBM); every call that hands the layout out restores it first (canonicalize).
  std::map<unsigned, unsigned> ref_pos_;
  void canonicalize();
  GateQueue queue_;
  std::vector<Gate> window_;
  std::mt19937 rng_;
  std::uniform_real_distribution<double> uni_{0.0, 1.0};
  hipStream_t stream_ = nullptr;
  double *d_partials_ = nullptr, *d_result_ = nullptr, *h_result_ = nullptr;
  int *d_flags_ = nullptr, *h_flags_ = nullptr;
  unsigned long long *d_key_ = nullptr, *h_key_ = nullptr;
  double *d_chunks_ = nullptr, *h_chunks_ = nullptr;
  size_t chunks_cap_ = 0;
  bool profiling_ = false;
  uint64_t stat_passes_ = 0, stat_timed_ = 0;
  double stat_bytes_ = 0.0, stat_ms_ = 0.0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_used_;
  std::vector<hipEvent_t> ev_free_;
  hipEvent_t take_event();
  void harvest_events();
  void count_pass(double bytes) {
    stat_passes_++;
    stat_bytes_ += bytes;
  }
  char *d_arena_ = nullptr, *h_arena_ = nullptr;
  size_t arena_off_ = 0;
  static constexpr size_t kArenaBytes = size_t(8) << 20;
  // one tile pass uploads four arrays, each rounded up to 256 bytes
  static_assert(kArenaBytes >= kPassUploadBudget + 4 * 256, "a tile pass program must fit the upload arena");
};

}  // namespace qs
