// Batched circuits: many independent small circuits (uneven qubit counts) simulated together,
// each from |0...0>, returning <psi|H|psi> per circuit.  All states live in one HBM buffer; every
// circuit is planned into tile passes; pass p of every circuit runs in one launch; the whole run
// (state init, passes, energies) is captured once into a hipGraph and replayed with a single
// graph launch.  The tile shape is picked by an autotuner that times candidate shapes on the
// batch itself; the shape only changes speed, never the circuit each state undergoes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "planner.hpp"

namespace qs {

struct BatchCircuit {
  unsigned n = 0;
  std::vector<PGate> gates;             // positions = qubit indices 0..n-1
  std::vector<dev::PauliTerm> terms;    // H = sum of terms, coefficient and phase folded in
};

struct BatchShape {
  unsigned tmax = 12, rb = 4, lmin = 5;
  bool relabel = false;  // tile stores relabel qubits (schedule_tiles); energies follow the layout
};

class Batch {
 public:
  Batch(bool c64, std::vector<BatchCircuit> circuits, bool autotune);
  ~Batch();
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;

  // one run; with `graph` a single hipGraph launch, else the same kernels launched eagerly.
  // energies[c] for every circuit; returns the GPU time of the run in ms.
  double run(bool graph, double *energies);
  const BatchShape &shape() const { return shape_; }
  unsigned launches() const { return (unsigned)launch_jobs_.size(); }

 private:
  void build(const BatchShape &shape);
  void release();
  void destroy();  // release() plus the stream, events and pinned results
  void enqueue();  // every kernel of one run on stream_

  bool c64_;
  std::vector<BatchCircuit> circ_;
  BatchShape shape_;
  bool two_qubit_ops_ = true;  // any 2-qubit register op in the batch (picks the tile kernel)
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  // device data
  void *psi_ = nullptr;
  uint64_t total_amps_ = 0;
  uint64_t *d_state_off_ = nullptr;
  void *d_stages_ = nullptr, *d_rops_ = nullptr, *d_tops_ = nullptr, *d_mats_ = nullptr;
  std::vector<dev::BatchJob *> d_jobs_;
  std::vector<uint32_t *> d_block_job_;
  std::vector<std::vector<dev::BatchJob>> launch_jobs_;
  std::vector<uint32_t> launch_blocks_, launch_tmax_;
  dev::EnergyJob *d_ejobs_ = nullptr;
  dev::EnergyBlock *d_eblock_ = nullptr;
  uint32_t e_blocks_ = 0;
  dev::BatchTerm *d_terms_ = nullptr;
  double *d_partials_ = nullptr, *d_energies_ = nullptr, *h_energies_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
};

}  // namespace qs
