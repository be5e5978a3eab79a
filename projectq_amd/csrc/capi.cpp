// extern "C" boundary of the simulator library: one entry point per method of the reference's
// pybind11 class (projectq/backends/_sim/_cppsim.cpp:43-67); declarations and conventions in
// include/projectq_amd.h.  Exceptions become status codes (the reference's pybind11 exception
// translation: runtime_error -> RuntimeError, length_error/invalid_argument -> ValueError,
// bad_alloc -> MemoryError).
#include "projectq_amd.h"

#include <cstring>
#include <string>

#include "batch.hpp"
#include "engine.hpp"
#include "mlp.hpp"
#include "noalloc.hpp"

struct qs_simulator {
  qs::Engine engine;
  int device;
  qs_simulator(uint32_t seed, int prec, int fusion, int planner, unsigned hint, int dev)
      : engine(seed, prec, fusion, planner, hint), device(dev) {}
};

namespace {

thread_local std::string g_last_error;

template <class F>
int guarded(qs_simulator *s, F &&f) {
  try {
    if (s) QS_HIP(hipSetDevice(s->device));
    f();
    return QS_OK;
  } catch (const qs::CallbackError &e) {
    g_last_error = e.what();
    return QS_ERR_CALLBACK;
  } catch (const qs::DeviceError &e) {
    g_last_error = e.what();
    return QS_ERR_DEVICE;
  } catch (const std::length_error &e) {
    g_last_error = e.what();
    return QS_ERR_VALUE;
  } catch (const std::invalid_argument &e) {
    g_last_error = e.what();
    return QS_ERR_VALUE;
  } catch (const std::bad_alloc &) {
    g_last_error = "out of device or host memory";
    return QS_ERR_MEMORY;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return QS_ERR_RUNTIME;
  } catch (...) {
    g_last_error = "unknown error";
    return QS_ERR_RUNTIME;
  }
}

std::vector<unsigned> ids_of(const uint32_t *p, size_t n) { return std::vector<unsigned>(p, p + n); }

std::vector<qs::PauliString> terms_of(const uint32_t *idx, const uint8_t *ops, const size_t *off, size_t nt) {
  std::vector<qs::PauliString> t(nt);
  for (size_t i = 0; i < nt; ++i)
    for (size_t j = off[i]; j < off[i + 1]; ++j) t[i].ops.push_back({idx[j], (char)ops[j]});
  return t;
}

std::vector<std::vector<unsigned>> regs_of(const uint32_t *q, const uint32_t *sz, size_t nr) {
  std::vector<std::vector<unsigned>> r(nr);
  size_t k = 0;
  for (size_t i = 0; i < nr; ++i)
    for (uint32_t j = 0; j < sz[i]; ++j) r[i].push_back(q[k++]);
  return r;
}

}  // namespace

#define API extern "C" __attribute__((visibility("default")))

API const char *qs_last_error(void) { return g_last_error.c_str(); }
API int qs_abi_version(void) { return 1; }

API int qs_create(uint32_t seed, int precision, int fusion_mode, int planner, uint32_t capacity_hint_qubits,
                  qs_simulator **out) {
  return guarded(nullptr, [&] {
    int dev = 0;
    QS_HIP(hipGetDevice(&dev));
    *out = new qs_simulator(seed, precision, fusion_mode, planner, capacity_hint_qubits, dev);
  });
}

API void qs_destroy(qs_simulator *s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  delete s;
}

API int qs_allocate_qubit(qs_simulator *s, uint32_t id) {
  return guarded(s, [&] { s->engine.allocate_qubit(id); });
}
API int qs_deallocate_qubit(qs_simulator *s, uint32_t id) {
  return guarded(s, [&] { s->engine.deallocate_qubit(id); });
}
API int qs_get_classical_value(qs_simulator *s, uint32_t id, double tol, int *value) {
  return guarded(s, [&] { *value = s->engine.get_classical_value(id, tol) ? 1 : 0; });
}
API int qs_is_classical(qs_simulator *s, uint32_t id, double tol, int *value) {
  return guarded(s, [&] { *value = s->engine.is_classical(id, tol) ? 1 : 0; });
}
API int qs_measure_qubits(qs_simulator *s, const uint32_t *ids, size_t n, uint8_t *results) {
  return guarded(s, [&] {
    auto r = s->engine.measure_qubits(ids_of(ids, n));
    std::memcpy(results, r.data(), r.size());
  });
}
API int qs_apply_controlled_gate(qs_simulator *s, const double *matrix, const uint32_t *ids, size_t n_ids,
                                 const uint32_t *ctrl, size_t n_ctrl) {
  return guarded(s, [&] {
    if (n_ids > 12) throw qs::ValueError("apply_controlled_gate: gate too large");
    qs::Gate g;
    const size_t d = size_t(1) << n_ids;
    g.m.resize(d * d);
    for (size_t i = 0; i < d * d; ++i) g.m[i] = qs::cplx(matrix[2 * i], matrix[2 * i + 1]);
    g.ids = ids_of(ids, n_ids);
    g.ctrl = ids_of(ctrl, n_ctrl);
    s->engine.apply_controlled_gate(std::move(g));
  });
}
API int qs_emulate_math(qs_simulator *s, qs_math_fn fn, void *user, const uint32_t *reg_qubits,
                        const uint32_t *reg_sizes, size_t n_regs, const uint32_t *ctrl, size_t n_ctrl) {
  return guarded(s, [&] {
    s->engine.emulate_math_callback([&](int32_t *v, size_t nr) { return fn(user, v, nr) == 0; },
                                    regs_of(reg_qubits, reg_sizes, n_regs), ids_of(ctrl, n_ctrl));
  });
}
API int qs_emulate_math_op(qs_simulator *s, int opcode, const int64_t *params, size_t n_params,
                           const uint32_t *reg_qubits, const uint32_t *reg_sizes, size_t n_regs,
                           const uint32_t *ctrl, size_t n_ctrl) {
  return guarded(s, [&] {
    if (opcode < QS_MATH_ADD_CONSTANT || opcode > QS_MATH_INVERSE_DIVIDE_QUANTUM)
      throw qs::ValueError("emulate_math: unknown opcode");
    std::vector<int64_t> prm(params, params + n_params);
    if (opcode <= QS_MATH_MULTIPLY_BY_CONSTANT_MOD_N) {
      size_t need = opcode == QS_MATH_ADD_CONSTANT ? 1 : 2;
      if (n_params < need) throw qs::ValueError("emulate_math: missing constant parameters");
      if (opcode != QS_MATH_ADD_CONSTANT && (int32_t)prm[1] == 0) throw qs::ValueError("emulate_math: N must be non-zero");
    } else {
      size_t need = (opcode == QS_MATH_ADD_QUANTUM || opcode == QS_MATH_INVERSE_ADD_QUANTUM ||
                     opcode == QS_MATH_SUBTRACT_QUANTUM) ? 2 : 3;
      if (n_regs < need || n_regs > 3) throw qs::ValueError("emulate_math: wrong number of registers for this function");
    }
    s->engine.emulate_math_op(opcode, prm, regs_of(reg_qubits, reg_sizes, n_regs), ids_of(ctrl, n_ctrl));
  });
}
API int qs_get_expectation_value(qs_simulator *s, const uint32_t *term_idx, const uint8_t *term_ops,
                                 const size_t *term_offsets, size_t n_terms, const double *coeffs,
                                 const uint32_t *ids, size_t n_ids, double *result) {
  return guarded(s, [&] {
    *result = s->engine.get_expectation_value(terms_of(term_idx, term_ops, term_offsets, n_terms),
                                              std::vector<double>(coeffs, coeffs + n_terms), ids_of(ids, n_ids));
  });
}
API int qs_expectation_value_noalloc(const void *device_state, uint32_t n_qubits, int precision,
                                    const uint32_t *term_idx, const uint8_t *term_ops, const size_t *term_offsets,
                                    size_t n_terms, const double *coeffs, const uint32_t *positions,
                                    size_t n_positions, double *result) {
  return guarded(nullptr, [&] {
    if (precision != QS_C128 && precision != QS_C64) throw qs::ValueError("precision must be QS_C128 or QS_C64");
    *result = qs::expectation_noalloc(device_state, n_qubits, precision == QS_C64, term_idx, term_ops, term_offsets,
                                      n_terms, coeffs, positions, n_positions);
  });
}
API int qs_apply_qubit_operator(qs_simulator *s, const uint32_t *term_idx, const uint8_t *term_ops,
                                const size_t *term_offsets, size_t n_terms, const double *coeffs,
                                const uint32_t *ids, size_t n_ids) {
  return guarded(s, [&] {
    std::vector<qs::cplx> c(n_terms);
    for (size_t i = 0; i < n_terms; ++i) c[i] = qs::cplx(coeffs[2 * i], coeffs[2 * i + 1]);
    s->engine.apply_qubit_operator(terms_of(term_idx, term_ops, term_offsets, n_terms), c, ids_of(ids, n_ids));
  });
}
API int qs_emulate_time_evolution(qs_simulator *s, const uint32_t *term_idx, const uint8_t *term_ops,
                                  const size_t *term_offsets, size_t n_terms, const double *coeffs, double time,
                                  const uint32_t *ids, size_t n_ids, const uint32_t *ctrl, size_t n_ctrl) {
  return guarded(s, [&] {
    s->engine.emulate_time_evolution(terms_of(term_idx, term_ops, term_offsets, n_terms),
                                     std::vector<double>(coeffs, coeffs + n_terms), time, ids_of(ids, n_ids),
                                     ids_of(ctrl, n_ctrl));
  });
}
API int qs_get_probability(qs_simulator *s, const uint8_t *bits, const uint32_t *ids, size_t n, double *result) {
  return guarded(s, [&] {
    *result = s->engine.get_probability(std::vector<uint8_t>(bits, bits + n), ids_of(ids, n));
  });
}
API int qs_get_amplitude(qs_simulator *s, const uint8_t *bits, const uint32_t *ids, size_t n, double *result) {
  return guarded(s, [&] {
    qs::cplx a = s->engine.get_amplitude(std::vector<uint8_t>(bits, bits + n), ids_of(ids, n));
    result[0] = a.real();
    result[1] = a.imag();
  });
}
API int qs_set_wavefunction(qs_simulator *s, const double *wf, size_t n_amps, const uint32_t *ordering,
                            size_t n_ids) {
  return guarded(s, [&] { s->engine.set_wavefunction(wf, n_amps, ids_of(ordering, n_ids)); });
}
API int qs_collapse_wavefunction(qs_simulator *s, const uint32_t *ids, size_t n_ids, const uint8_t *values,
                                 size_t n_values) {
  return guarded(s, [&] {
    s->engine.collapse_wavefunction(ids_of(ids, n_ids), std::vector<uint8_t>(values, values + n_values));
  });
}
API int qs_run(qs_simulator *s) {
  return guarded(s, [&] { s->engine.run(); });
}
API int qs_cheat_sizes(qs_simulator *s, size_t *n_qubits, size_t *n_amps) {
  return guarded(s, [&] { s->engine.cheat_sizes(n_qubits, n_amps); });
}
API int qs_cheat(qs_simulator *s, uint32_t *ids, uint32_t *positions, double *state) {
  return guarded(s, [&] { s->engine.cheat(ids, positions, state); });
}

API int qs_device_state(qs_simulator *s, void **device_ptr, uint64_t *n_amps, int *precision) {
  return guarded(s, [&] { s->engine.device_state(device_ptr, n_amps, precision); });
}
API int qs_qubit_map(qs_simulator *s, uint32_t *ids, uint32_t *positions, size_t capacity, size_t *n_qubits) {
  return guarded(s, [&] { s->engine.qubit_map(ids, positions, capacity, n_qubits); });
}
API int qs_save_state(qs_simulator *s, const char *path) {
  return guarded(s, [&] {
    if (!path) throw qs::ValueError("save_state: no path");
    s->engine.save_state(path);
  });
}
API int qs_load_state(qs_simulator *s, const char *path) {
  return guarded(s, [&] {
    if (!path) throw qs::ValueError("load_state: no path");
    s->engine.load_state(path);
  });
}
API int qs_set_wavefunction_device(qs_simulator *s, const void *device_ptr, uint64_t n_amps, int precision,
                                   const uint32_t *ordering, size_t n_ordering) {
  return guarded(s, [&] {
    if (!device_ptr && n_amps) throw qs::ValueError("set_wavefunction_device: null device pointer");
    std::vector<unsigned> ord(ordering, ordering + n_ordering);
    s->engine.set_wavefunction_device(device_ptr, n_amps, precision, ord);
  });
}

API int qs_synchronize(qs_simulator *s) {
  return guarded(s, [&] { s->engine.synchronize(); });
}
API int qs_set_profiling(qs_simulator *s, int enabled) {
  return guarded(s, [&] { s->engine.set_profiling(enabled != 0); });
}
API int qs_get_stats(qs_simulator *s, uint64_t *passes, double *bytes, double *gpu_ms, uint64_t *timed) {
  return guarded(s, [&] { s->engine.get_stats(passes, bytes, gpu_ms, timed); });
}
API int qs_reset_stats(qs_simulator *s) {
  return guarded(s, [&] { s->engine.reset_stats(); });
}
API int qs_overlap(qs_simulator *a, qs_simulator *b, double *result) {
  return guarded(a, [&] {
    if (a->device != b->device) throw qs::ValueError("overlap: the simulators live on different GPUs");
    qs::cplx v = a->engine.overlap(b->engine);
    result[0] = v.real();
    result[1] = v.imag();
  });
}

API int qs_compare_states(qs_simulator *a, qs_simulator *b, double *max_abs_diff, uint64_t *n_unequal) {
  return guarded(a, [&] {
    if (a->device != b->device) throw qs::ValueError("compare: the simulators live on different GPUs");
    a->engine.compare(b->engine, max_abs_diff, n_unequal);
  });
}

struct qs_batch {
  qs::Batch batch;
  int device;
  qs_batch(bool c64, std::vector<qs::BatchCircuit> c, bool tune, int dev) : batch(c64, std::move(c), tune), device(dev) {}
};

API int qs_batch_create(int precision, size_t n_circuits, const uint32_t *n_qubits, const size_t *gate_offsets,
                        const uint32_t *gate_k, const uint32_t *gate_nctrl, const uint32_t *gate_qubits,
                        const double *matrices, const size_t *term_offsets, const size_t *op_offsets,
                        const uint32_t *op_qubits, const uint8_t *op_kinds, const double *coeffs, int autotune,
                        qs_batch **out) {
  return guarded(nullptr, [&] {
    if (precision != QS_C128 && precision != QS_C64) throw qs::ValueError("batch: unknown precision");
    std::vector<qs::BatchCircuit> circ(n_circuits);
    size_t qi = 0, mi = 0;
    for (size_t b = 0; b < n_circuits; ++b) {
      circ[b].n = n_qubits[b];
      if (circ[b].n < 1 || circ[b].n > 30) throw qs::ValueError("batch: circuits hold 1..30 qubits");
      for (size_t g = gate_offsets[b]; g < gate_offsets[b + 1]; ++g) {
        qs::PGate pg;
        const uint32_t k = gate_k[g];
        if (k < 1 || k > 5) throw qs::ValueError("batch: gates act on 1..5 qubits");
        for (uint32_t l = 0; l < k; ++l) pg.tpos.push_back(gate_qubits[qi++]);
        for (uint32_t l = 0; l < gate_nctrl[g]; ++l) {
          const uint32_t c = gate_qubits[qi++];
          if (c >= 64) throw qs::ValueError("batch: control outside its circuit");
          pg.cmask |= uint64_t(1) << c;
        }
        const size_t d = size_t(1) << k;
        pg.m.resize(d * d);
        for (size_t e = 0; e < d * d; ++e, ++mi) pg.m[e] = qs::cplx(matrices[2 * mi], matrices[2 * mi + 1]);
        circ[b].gates.push_back(std::move(pg));
      }
      for (size_t t = term_offsets[b]; t < term_offsets[b + 1]; ++t) {
        std::vector<std::pair<uint32_t, char>> ops;
        for (size_t o = op_offsets[t]; o < op_offsets[t + 1]; ++o) {
          if (op_qubits[o] >= circ[b].n) throw qs::ValueError("batch: Pauli string outside its circuit");
          ops.push_back({op_qubits[o], (char)op_kinds[o]});
        }
        qs::dev::PauliTerm pt{};
        qs::cplx w;
        qs::encode_pauli(ops, &pt.x, &pt.z, &w);
        // the energy kernel pairs i with i ^ x, which needs a Hermitian string; an anti-Hermitian
        // one (a qubit named twice) has Re<psi|P|psi> = 0 and is left out
        if (!qs::dev::pauli_hermitian(pt.x, pt.z, w)) continue;
        const qs::cplx c = coeffs[t] * w;
        pt.c_re = c.real();
        pt.c_im = c.imag();
        circ[b].terms.push_back(pt);
      }
    }
    int dev = 0;
    QS_HIP(hipGetDevice(&dev));
    *out = new qs_batch(precision == QS_C64, std::move(circ), autotune != 0, dev);
  });
}

API int qs_batch_run(qs_batch *b, int use_graph, double *energies, double *gpu_ms) {
  return guarded(nullptr, [&] {
    QS_HIP(hipSetDevice(b->device));
    double ms = b->batch.run(use_graph != 0, energies);
    if (gpu_ms) *gpu_ms = ms;
  });
}

API int qs_batch_shape(qs_batch *b, uint32_t *t, uint32_t *rb, uint32_t *lmin, uint32_t *launches) {
  return guarded(nullptr, [&] {
    *t = b->batch.shape().tmax;
    *rb = b->batch.shape().rb;
    *lmin = b->batch.shape().lmin;
    *launches = b->batch.launches();
  });
}

API int qs_batch_relabel(qs_batch *b, int *relabel) {
  return guarded(nullptr, [&] { *relabel = b->batch.shape().relabel ? 1 : 0; });
}

API void qs_batch_destroy(qs_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  delete b;
}

API int qs_set_tile_params(qs_simulator *s, uint32_t t, uint32_t rb, uint32_t low) {
  return guarded(s, [&] { s->engine.set_tile_params(t, rb, low); });
}
API int qs_set_slot_rule(qs_simulator *s, int rule) {
  return guarded(s, [&] { s->engine.set_slot_rule(rule); });
}
API int qs_set_relabel(qs_simulator *s, int mode) {
  return guarded(s, [&] { s->engine.set_relabel(mode); });
}
API int qs_last_plan_relabels(qs_simulator *s, uint32_t *passes) {
  return guarded(s, [&] { *passes = s->engine.last_plan().relabels; });
}
API int qs_last_plan_rule(qs_simulator *s, int *rule) {
  return guarded(s, [&] { *rule = s->engine.last_plan().rule; });
}
API int qs_set_recording(qs_simulator *s, int on) {
  return guarded(s, [&] { s->engine.set_recording(on != 0); });
}
API int qs_take_records(qs_simulator *s, float *features, float *ms, size_t cap, size_t *rows) {
  return guarded(s, [&] { *rows = s->engine.take_records(features, ms, cap); });
}
API int qs_set_planner_model(qs_simulator *s, const float *w, size_t n) {
  return guarded(s, [&] {
    if (n != 0 && n != (size_t)qs::mlp::kWeights) throw qs::ValueError("planner model: wrong number of weights");
    s->engine.set_model(std::vector<float>(w, w + n));
  });
}
API int qs_last_plan(qs_simulator *s, int *kind, uint32_t *t, uint32_t *rb, uint32_t *low, uint32_t *passes,
                     uint32_t *cands, double *pred) {
  return guarded(s, [&] {
    const auto &p = s->engine.last_plan();
    *kind = p.kind;
    *t = p.tmax;
    *rb = p.rb;
    *low = p.lmin;
    *passes = p.passes;
    *cands = p.candidates;
    *pred = p.predicted_ms;
  });
}
API size_t qs_mlp_num_weights(void) { return (size_t)qs::mlp::kWeights; }
API int qs_mlp_predict(const float *w, const float *X, size_t rows, float *out) {
  return guarded(nullptr, [&] { qs::mlp::predict(w, X, rows, out); });
}
API int qs_mlp_loss_grad(const float *w, const float *X, const float *T, size_t rows, float *loss, float *grad) {
  return guarded(nullptr, [&] { qs::mlp::loss_grad(w, X, T, rows, loss, grad); });
}
API int qs_mlp_train(const float *X, const float *T, size_t rows, int epochs, int batch, float lr, uint32_t seed,
                     const float *w_init, float *w_out, float *final_loss) {
  return guarded(nullptr, [&] { *final_loss = qs::mlp::train(X, T, rows, epochs, batch, lr, seed, w_init, w_out); });
}
