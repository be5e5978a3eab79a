/*
 * projectq_amd.h -- C ABI of the MI355X state-vector simulator.
 *
 * One opaque handle per simulator.  Every entry point mirrors one method of the pybind11 class
 * `projectq.backends._sim._cppsim.Simulator` (projectq/backends/_sim/_cppsim.cpp:43-67), whose
 * implementation lives in projectq/backends/_sim/_cppkernels/simulator.hpp.  The Python binding
 * (projectq_amd/_cppsim.py) keeps that class's method names and argument conventions, so
 * `projectq.backends.Simulator()._simulator = projectq_amd.Simulator(seed)` is a drop-in.
 *
 * Conventions shared by all entry points
 *  - Qubit ids are uint32.  A gate matrix of k target qubits is 2^k x 2^k complex, passed
 *    row-major as interleaved (re, im) doubles: m[2*(r*2^k + c)] = Re M[r][c].  Bit l of the
 *    row/column index belongs to ids[l] (simulator.hpp:487-527 with fusion.hpp:65-109).
 *  - Register layouts (emulate_math) are passed as all qubit ids concatenated plus one size per
 *    register; register r's qubit j is bit j of the register value (little endian).
 *  - Pauli terms (expectation value, qubit operator, time evolution) are passed as
 *    concatenated (local index, op) pairs plus offsets[n_terms + 1]; op is 'X', 'Y' or 'Z' and
 *    local index l refers to ids[l] (simulator.hpp:538-550).
 *  - Bits / booleans are uint8 (0 or 1).
 *  - Every function returns QS_OK or an error class; qs_last_error() holds the message of the
 *    last failure on the calling thread.  The classes map to the C++ exception types the
 *    reference throws: runtime_error -> QS_ERR_RUNTIME, length_error/invalid_argument ->
 *    QS_ERR_VALUE, bad_alloc -> QS_ERR_MEMORY.
 *
 * Thread safety: a handle must not be used from two threads at once (same as the reference,
 * which is called under the GIL).  Different handles are independent: unlike the reference's
 * process-wide static scratch buffers (simulator.hpp:574-578), every handle owns its buffers.
 */
#ifndef PROJECTQ_AMD_H_
#define PROJECTQ_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qs_simulator qs_simulator;

/* status codes */
#define QS_OK 0
#define QS_ERR_RUNTIME 1  /* std::runtime_error   -> RuntimeError */
#define QS_ERR_VALUE 2    /* std::length_error / std::invalid_argument -> ValueError */
#define QS_ERR_MEMORY 3   /* std::bad_alloc / hipErrorOutOfMemory -> MemoryError */
#define QS_ERR_DEVICE 4   /* any other HIP failure -> RuntimeError */
#define QS_ERR_CALLBACK 5 /* a user callback (emulate_math) reported failure */

/* precision of the resident state */
#define QS_C128 0 /* complex<double>, the reference's calc_type (simulator.hpp:39-40) */
#define QS_C64 1  /* complex<float> state; norms, probabilities and sums stay FP64 */

/* how apply_controlled_gate queues gates */
#define QS_FUSION_REFERENCE 0 /* greedy min-4/max-5 rule of simulator.hpp:204-222, run() drains */
#define QS_FUSION_WINDOW 1    /* queue every gate until run(); the planner partitions the window */

/* planner used for QS_FUSION_WINDOW */
#define QS_PLANNER_GREEDY 0   /* greedy max-5 blocks (fusion.hpp restated over the whole window) */
#define QS_PLANNER_LEARNED 1  /* tile passes whose tiles an MLP cost model picks; QS_PLANNER_TILE's
                                 heuristic is its fallback */
#define QS_PLANNER_TILE 2     /* tile passes: every gate whose targets lie in a 2^t-amplitude LDS
                                 tile is applied during one streaming pass (DAG order kept) */

/* emulate_math opcodes (qs_emulate_math_op).  The first three are the reference's native fast
 * paths (simulator.hpp:271-290); the rest are the integer functions of projectq/libs/math/_gates.py,
 * which the reference evaluates through a Python callback. */
#define QS_MATH_ADD_CONSTANT 0            /* x += a                      (_gates.py:20-62)  */
#define QS_MATH_ADD_CONSTANT_MOD_N 1      /* x = (x + a) % N             (_gates.py:82-133) */
#define QS_MATH_MULTIPLY_BY_CONSTANT_MOD_N 2 /* x = (x * a) % N          (_gates.py:164-213) */
#define QS_MATH_ADD_QUANTUM 3             /* (a, b[, c]): b += a, wrap + flip carry (_gates.py:251-265) */
#define QS_MATH_SUBTRACT_QUANTUM 4        /* (a, b): b -= a              (_gates.py:298-344) */
#define QS_MATH_INVERSE_ADD_QUANTUM 5     /* (a, b[, c]): flip c, b -= a (_gates.py:268-296) */
#define QS_MATH_COMPARATOR 6              /* (a, b, c): c ^= (b < a)     (_gates.py:347-400) */
#define QS_MATH_MULTIPLY_QUANTUM 7        /* (a, b, c): c += a*b         (_gates.py:488-538) */
#define QS_MATH_INVERSE_MULTIPLY_QUANTUM 8 /* (a, b, c): c -= a*b        (_gates.py:541-570) */
#define QS_MATH_DIVIDE_QUANTUM 9          /* (dividend, remainder, divisor) (_gates.py:403-464) */
#define QS_MATH_INVERSE_DIVIDE_QUANTUM 10 /* (remainder, quotient, divisor) (_gates.py:467-485) */

/* callback for the generic emulate_math: rewrite values[0..n_regs) in place; return 0 on
 * success, non-zero to abort the call with QS_ERR_CALLBACK. */
typedef int (*qs_math_fn)(void *user, int32_t *values, size_t n_regs);

const char *qs_last_error(void);
int qs_abi_version(void);

/* Simulator(seed) -- simulator.hpp:48-53.  capacity_hint_qubits reserves HBM for that many
 * qubits up front (0 = grow on demand). */
int qs_create(uint32_t seed, int precision, int fusion_mode, int planner,
              uint32_t capacity_hint_qubits, qs_simulator **out);
void qs_destroy(qs_simulator *sim);

/* simulator.hpp:55-74 */
int qs_allocate_qubit(qs_simulator *sim, uint32_t id);
/* simulator.hpp:194-202 (is_classical check, then collapse_vector with shrink) */
int qs_deallocate_qubit(qs_simulator *sim, uint32_t id);
/* simulator.hpp:76-91 */
int qs_get_classical_value(qs_simulator *sim, uint32_t id, double tol, int *value);
/* simulator.hpp:93-108 */
int qs_is_classical(qs_simulator *sim, uint32_t id, double tol, int *value);
/* simulator.hpp:145-192 */
int qs_measure_qubits(qs_simulator *sim, const uint32_t *ids, size_t n_ids, uint8_t *results);
/* simulator.hpp:204-222 (+ fusion.hpp:56-63) */
int qs_apply_controlled_gate(qs_simulator *sim, const double *matrix, const uint32_t *ids,
                             size_t n_ids, const uint32_t *ctrl, size_t n_ctrl);
/* simulator.hpp:224-269 with a host callback; the callback sees each distinct tuple of
 * register values once. */
int qs_emulate_math(qs_simulator *sim, qs_math_fn fn, void *user, const uint32_t *reg_qubits,
                    const uint32_t *reg_sizes, size_t n_regs, const uint32_t *ctrl,
                    size_t n_ctrl);
/* simulator.hpp:224-290: integer functions evaluated natively (opcodes above); params holds
 * a then N for the constant forms and is ignored otherwise. */
int qs_emulate_math_op(qs_simulator *sim, int opcode, const int64_t *params, size_t n_params,
                       const uint32_t *reg_qubits, const uint32_t *reg_sizes, size_t n_regs,
                       const uint32_t *ctrl, size_t n_ctrl);
/* simulator.hpp:292-322 -> sum_t coeff_t * Re<psi|P_t|psi> */
int qs_get_expectation_value(qs_simulator *sim, const uint32_t *term_idx, const uint8_t *term_ops,
                             const size_t *term_offsets, size_t n_terms, const double *coeffs,
                             const uint32_t *ids, size_t n_ids, double *result);
/* No counterpart in the reference: qs_get_expectation_value for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and the same result as qs_get_expectation_value, on a caller-owned state of 2^n_qubits
 * amplitudes at `device_state` on the current GPU (complex128 for QS_C128, complex64 for QS_C64,
 * e.g. from qs_device_state); local index l of a string is the qubit at bit position
 * positions[l].  Allocates no device or host memory: the strings, partial sums and per-string
 * results use a fixed 224 KiB device array per GPU (held until the process ends).  At most 4096
 * strings per call (QS_ERR_VALUE beyond).  Synchronous on the null stream; calls that overlap on
 * one GPU share that array and race. */
int qs_expectation_value_noalloc(const void *device_state, uint32_t n_qubits, int precision,
                                 const uint32_t *term_idx, const uint8_t *term_ops, const size_t *term_offsets,
                                 size_t n_terms, const double *coeffs, const uint32_t *positions,
                                 size_t n_positions, double *result);
/* simulator.hpp:324-350; coeffs are complex (interleaved) */
int qs_apply_qubit_operator(qs_simulator *sim, const uint32_t *term_idx, const uint8_t *term_ops,
                            const size_t *term_offsets, size_t n_terms, const double *coeffs,
                            const uint32_t *ids, size_t n_ids);
/* simulator.hpp:386-438; coeffs are real */
int qs_emulate_time_evolution(qs_simulator *sim, const uint32_t *term_idx,
                              const uint8_t *term_ops, const size_t *term_offsets,
                              size_t n_terms, const double *coeffs, double time,
                              const uint32_t *ids, size_t n_ids, const uint32_t *ctrl,
                              size_t n_ctrl);
/* simulator.hpp:352-368 */
int qs_get_probability(qs_simulator *sim, const uint8_t *bits, const uint32_t *ids, size_t n_ids,
                       double *result);
/* simulator.hpp:370-384; result = (re, im) */
int qs_get_amplitude(qs_simulator *sim, const uint8_t *bits, const uint32_t *ids, size_t n_ids,
                     double *result);
/* simulator.hpp:440-454; wf = 2^n interleaved complex */
int qs_set_wavefunction(qs_simulator *sim, const double *wf, size_t n_amps,
                        const uint32_t *ordering, size_t n_ids);
/* simulator.hpp:456-485 */
int qs_collapse_wavefunction(qs_simulator *sim, const uint32_t *ids, size_t n_ids,
                             const uint8_t *values, size_t n_values);
/* simulator.hpp:487-527 */
int qs_run(qs_simulator *sim);
/* simulator.hpp:529-532, split in two calls: sizes, then copies into caller-owned buffers.
 * qs_cheat_sizes runs pending gates; state is 2^n interleaved complex128. */
int qs_cheat_sizes(qs_simulator *sim, size_t *n_qubits, size_t *n_amps);
int qs_cheat(qs_simulator *sim, uint32_t *ids, uint32_t *positions, double *state);

/* ---- state access without host copies, checkpoint / restore (no counterpart in the reference;
 * they extend cheat() and set_wavefunction() for states too large for a host list) ---------- */

/* Run pending gates and return the device address of the state: n_amps interleaved amplitudes,
 * complex128 (*precision == QS_C128) or complex64 (QS_C64), index bit p <-> the qubit at position
 * p (the map qs_cheat reports).  Valid until the next call on this simulator that applies gates,
 * allocates, deallocates or measures (those may move the state). */
int qs_device_state(qs_simulator *sim, void **device_ptr, uint64_t *n_amps, int *precision);
/* The qubit map alone (no state copy): up to `capacity` (id, position) pairs, *n_qubits = all. */
int qs_qubit_map(qs_simulator *sim, uint32_t *ids, uint32_t *positions, size_t capacity, size_t *n_qubits);
/* Write the qubit map and the amplitudes to `path` (binary, little-endian: 32-byte header
 * "QSAMDST1", version 1, precision, qubit count, amplitude count; then (id, position) uint32
 * pairs; then the amplitudes in the simulator's precision), streamed through pinned memory. */
int qs_save_state(qs_simulator *sim, const char *path);
/* Restore a file written by qs_save_state.  A simulator without qubits allocates the file's
 * qubits; one with qubits must hold exactly the file's qubit ids.  A file of the other precision
 * is converted (complex128 -> complex64 rounds to nearest).  ValueError on a malformed file. */
int qs_load_state(qs_simulator *sim, const char *path);
/* set_wavefunction (simulator.hpp:440-454) from device memory: `device_ptr` holds n_amps
 * interleaved amplitudes of precision `precision` on this simulator's GPU (converted if needed);
 * ordering as in qs_set_wavefunction. */
int qs_set_wavefunction_device(qs_simulator *sim, const void *device_ptr, uint64_t n_amps, int precision,
                               const uint32_t *ordering, size_t n_ordering);

/* ---- extensions (no counterpart in the reference) ------------------------------------------
 * Used by measurement harnesses and by tests that compare precisions on one GPU. */

/* Block until every queued device operation of this simulator has finished. */
int qs_synchronize(qs_simulator *sim);
/* Record HIP events around every gate pass (enabled != 0) so qs_get_stats can report the GPU
 * time of the passes; off by default. */
int qs_set_profiling(qs_simulator *sim, int enabled);
/* Counters since creation or the last qs_reset_stats: gate passes launched, their algorithmic
 * HBM bytes (2 * bytes per amplitude * amplitudes touched), and -- for passes launched with
 * profiling on -- the summed GPU time in ms and how many passes it covers.  Synchronises. */
int qs_get_stats(qs_simulator *sim, uint64_t *passes, double *bytes, double *gpu_ms,
                 uint64_t *timed_passes);
int qs_reset_stats(qs_simulator *sim);
/* <a|b> over two simulators on the same GPU holding the same qubits at the same positions
 * (precisions may differ); accumulated in FP64.  result = (re, im). */
int qs_overlap(qs_simulator *a, qs_simulator *b, double *result);
/* Element-wise comparison of the states of two simulators on the same GPU holding the same qubits
 * at the same positions (precisions may differ): *max_abs_diff = max_i |a_i - b_i| (FP64) and
 * *n_unequal = the number of amplitudes whose stored values differ. */
int qs_compare_states(qs_simulator *a, qs_simulator *b, double *max_abs_diff, uint64_t *n_unequal);

/* ---- planner controls and the learned cost model (extensions, no counterpart in the reference) --
 * Tile shape of QS_PLANNER_TILE: tile bits (6..13; 12 for complex128), register bits (4 or 5) and
 * low contiguous bits every tile keeps. */
int qs_set_tile_params(qs_simulator *sim, uint32_t tile_bits, uint32_t reg_bits, uint32_t low_bits);
/* Rule for the free tile slots of QS_PLANNER_TILE (the groupings its passes get): */
#define QS_SLOT_LOOKAHEAD 0 /* highest decayed lookahead score (default) */
#define QS_SLOT_EARLIEST 1  /* positions the earliest unscheduled gates need, in order of need */
int qs_set_slot_rule(qs_simulator *sim, int rule);
/* Qubit relabelling at tile stores (no counterpart in the reference): each tile pass moves into
   the low positions, which every tile holds, the qubits of its tile that the remaining gates need
   first; the exchange costs nothing (the store writes the tile with the bits exchanged) and the
   qubit map follows.  -1: default (on for QS_PLANNER_LEARNED, off for QS_PLANNER_TILE), 0: off,
   1: on.  Results are the same up to rounding; only the pass count changes. */
int qs_set_relabel(qs_simulator *sim, int mode);
/* How many passes of the last window's tile plan relabelled qubits at their stores. */
int qs_last_plan_relabels(qs_simulator *sim, uint32_t *passes);
/* The slot rule of the last window's tile plan (the learned planner tries both). */
int qs_last_plan_rule(qs_simulator *sim, int *rule);
/* Record every window pass: QS_PASS_FEATURES features and its GPU time. */
#define QS_PASS_FEATURES 24
int qs_set_recording(qs_simulator *sim, int enabled);
/* Move up to `cap` records out (features row-major, ms); *rows = how many.  Synchronises. */
int qs_take_records(qs_simulator *sim, float *features, float *ms, size_t cap, size_t *rows);
/* Load (n_weights = qs_mlp_num_weights()) or clear (n_weights = 0) the model QS_PLANNER_LEARNED
 * uses; without a model the learned planner falls back to greedy max-5 blocks. */
int qs_set_planner_model(qs_simulator *sim, const float *weights, size_t n_weights);
/* The last window's plan: kind (0 greedy blocks, 1 tile passes), tile shape, passes, number of
 * candidate plans compared and the model's predicted time (ms). */
int qs_last_plan(qs_simulator *sim, int *kind, uint32_t *tile_bits, uint32_t *reg_bits,
                 uint32_t *low_bits, uint32_t *passes, uint32_t *candidates, double *predicted_ms);
/* The MLP (24 -> 64 -> 64 -> 1, ReLU, fp32) on the GPU: weight count, inference, loss and
 * gradient (mean squared error), and Adam training on minibatches (w_init may be NULL). */
size_t qs_mlp_num_weights(void);
int qs_mlp_predict(const float *weights, const float *features, size_t rows, float *out);
int qs_mlp_loss_grad(const float *weights, const float *features, const float *targets, size_t rows,
                     float *loss, float *grad);
int qs_mlp_train(const float *features, const float *targets, size_t rows, int epochs, int batch,
                 float lr, uint32_t seed, const float *w_init, float *w_out, float *final_loss);

/* ---- batched circuits (extension, no counterpart in the reference) ---------------------------
 * B independent circuits of uneven qubit counts, each on its own register starting in |0..0>,
 * run together on one GPU and reduced to <psi_b|H_b|psi_b>.  Used by variational workloads.
 * Circuit b owns gates [gate_offsets[b], gate_offsets[b+1]); gate g has gate_k[g] targets then
 * gate_nctrl[g] controls, listed in gate_qubits (qubit indices 0..n_b-1), and a 2^k x 2^k matrix
 * (interleaved complex, concatenated).  Circuit b's operator H_b owns terms
 * [term_offsets[b], term_offsets[b+1]); term t has ops [op_offsets[t], op_offsets[t+1]) of
 * (op_qubits, op_kinds 'X'|'Y'|'Z') and a real coefficient.  autotune != 0 times candidate tile
 * shapes on the batch and keeps the fastest. */
typedef struct qs_batch qs_batch;
int qs_batch_create(int precision, size_t n_circuits, const uint32_t *n_qubits,
                    const size_t *gate_offsets, const uint32_t *gate_k, const uint32_t *gate_nctrl,
                    const uint32_t *gate_qubits, const double *matrices,
                    const size_t *term_offsets, const size_t *op_offsets,
                    const uint32_t *op_qubits, const uint8_t *op_kinds, const double *coeffs,
                    int autotune, qs_batch **out);
/* One run: a single hipGraph launch (use_graph != 0) or the same kernels launched one by one.
 * energies[b] receives <psi_b|H_b|psi_b>; *gpu_ms the GPU time of the run. */
int qs_batch_run(qs_batch *batch, int use_graph, double *energies, double *gpu_ms);
/* Tile shape in use (after autotuning) and the number of pass launches per run. */
int qs_batch_shape(qs_batch *batch, uint32_t *tile_bits, uint32_t *reg_bits, uint32_t *low_bits,
                   uint32_t *launches);
/* Whether the chosen shape relabels qubits at its tile stores (the energies follow the layout). */
int qs_batch_relabel(qs_batch *batch, int *relabel);
void qs_batch_destroy(qs_batch *batch);

#ifdef __cplusplus
}
#endif

#endif /* PROJECTQ_AMD_H_ */
