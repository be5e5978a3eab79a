// Batched circuits (see batch.hpp).
#include <cstdlib>
#include "batch.hpp"

#include <algorithm>
#include <complex>
#include <cstring>
#include <limits>

namespace qs {

namespace {

template <class T>
T *to_device(const std::vector<T> &v) {
  if (v.empty()) return nullptr;
  T *d = nullptr;
  QS_HIP(hipMalloc(&d, v.size() * sizeof(T)));
  QS_HIP(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

std::vector<BatchShape> candidate_shapes(bool c64) {
  if (c64)
    return {{12, 4, 5, false}, {11, 4, 5, false}, {10, 4, 4, false}, {12, 5, 5, false},
            {10, 5, 5, false}, {12, 4, 4, true}, {11, 4, 4, true}, {10, 4, 4, true}, {12, 4, 5, true}};
  return {{12, 4, 4, false}, {11, 4, 4, false}, {10, 4, 4, false}, {12, 4, 4, true}};
}

}  // namespace

Batch::Batch(bool c64, std::vector<BatchCircuit> circuits, bool autotune) : c64_(c64), circ_(std::move(circuits)) {
  if (circ_.empty()) throw ValueError("batch: no circuits");
  for (auto &c : circ_) {
    if (c.n < 1 || c.n > 30) throw ValueError("batch: circuits hold 1..30 qubits");
    for (auto &g : c.gates) {
      if (g.tpos.empty() || g.tpos.size() > 5) throw ValueError("batch: gates act on 1..5 qubits");
      for (uint32_t p : g.tpos)
        if (p >= c.n) throw ValueError("batch: gate target outside its circuit");
      if (g.cmask >> c.n) throw ValueError("batch: control outside its circuit");
    }
    for (auto &t : c.terms)
      if ((t.x | t.z) >> c.n) throw ValueError("batch: Pauli string outside its circuit");
  }
  std::vector<BatchShape> shapes = candidate_shapes(c64_);
  long forced = -1;  // PROJECTQ_AMD_BATCH_SHAPE=k: candidate k only, no timing (tests)
  if (const char *env = std::getenv("PROJECTQ_AMD_BATCH_SHAPE")) {
    forced = std::strtol(env, nullptr, 10);
    if (forced < 0 || forced >= (long)shapes.size())
      throw ValueError("PROJECTQ_AMD_BATCH_SHAPE: no such candidate shape");
  }
  try {  // a constructor that throws runs no destructor: free what was made so far
    dev::prepare_tile_kernels();
    QS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    QS_HIP(hipEventCreate(&ev0_));
    QS_HIP(hipEventCreate(&ev1_));
    QS_HIP(hipHostMalloc((void **)&h_energies_, circ_.size() * sizeof(double), hipHostMallocDefault));
    if (forced >= 0 || !autotune) {
      build(shapes[forced >= 0 ? (size_t)forced : 0]);
      return;
    }
    // autotuner: time every candidate shape on this batch (graph launches), keep the fastest
    std::vector<double> e(circ_.size());
    double best = std::numeric_limits<double>::infinity();
    BatchShape best_shape = shapes[0];
    for (const BatchShape &s : shapes) {
      build(s);
      run(true, e.data());  // warm-up
      double ms = 0.0;
      for (int r = 0; r < 3; ++r) ms += run(true, e.data());
      if (ms < best) {
        best = ms;
        best_shape = s;
      }
    }
    build(best_shape);
  } catch (...) {
    destroy();
    throw;
  }
}

Batch::~Batch() { destroy(); }

void Batch::destroy() {
  release();
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (h_energies_) (void)hipHostFree(h_energies_);
  if (stream_) (void)hipStreamDestroy(stream_);
  ev0_ = ev1_ = nullptr;
  h_energies_ = nullptr;
  stream_ = nullptr;
}

void Batch::release() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
  for (auto p : d_jobs_) (void)hipFree(p);
  for (auto p : d_block_job_) (void)hipFree(p);
  d_jobs_.clear();
  d_block_job_.clear();
  launch_jobs_.clear();
  launch_blocks_.clear();
  launch_tmax_.clear();
  void *bufs[] = {psi_, d_state_off_, d_stages_, d_rops_, d_tops_, d_mats_, d_ejobs_, d_eblock_, d_terms_, d_partials_, d_energies_};
  for (void *b : bufs) (void)hipFree(b);
  psi_ = d_state_off_ = nullptr;
  d_stages_ = d_rops_ = d_tops_ = d_mats_ = nullptr;
  d_ejobs_ = nullptr;
  d_eblock_ = nullptr;
  d_terms_ = nullptr;
  d_partials_ = d_energies_ = nullptr;
}

void Batch::build(const BatchShape &shape) {
  release();
  shape_ = shape;
  // ---- plan every circuit; collect programs and per-pass jobs
  std::vector<dev::TileStage> stages;
  std::vector<dev::RegOp> rops;
  std::vector<dev::TileOp> tops;
  std::vector<std::complex<double>> mats;
  std::vector<uint64_t> state_off(circ_.size());
  std::vector<std::vector<dev::BatchJob>> jobs;  // by pass index
  std::vector<std::vector<size_t>> job_circuit;  // circuit of every job, by pass index
  std::vector<std::vector<uint32_t>> final_pos(circ_.size());  // label -> position after the passes
  total_amps_ = 0;
  for (size_t c = 0; c < circ_.size(); ++c) {
    const BatchCircuit &bc = circ_[c];
    state_off[c] = total_amps_;
    total_amps_ += uint64_t(1) << bc.n;
    std::vector<PGate> gates;
    for (const PGate &g : bc.gates)
      if (!(g.tmask() & g.cmask)) gates.push_back(g);
    gates = merge_single_qubit_runs(gates);
    const unsigned t = std::min(bc.n, shape.tmax);
    TileShape ts{bc.n, t, shape.lmin};
    std::vector<PPass> passes = schedule_tiles(gates, ts, nullptr, 256, 16.0, shape.relabel);
    std::vector<uint32_t> &phys = final_pos[c];
    std::vector<uint32_t> label(bc.n);
    phys.resize(bc.n);
    for (uint32_t q = 0; q < bc.n; ++q) phys[q] = label[q] = q;
    for (size_t p = 0; p < passes.size(); ++p) {
      TileProgram prog;
      build_tile_pass(gates, passes[p], t, shape.rb, c64_, prog);
      for (const auto &sw : prog.swaps) {
        const uint32_t a = label[sw.first], b = label[sw.second];
        std::swap(label[sw.first], label[sw.second]);
        phys[a] = sw.second;
        phys[b] = sw.first;
      }
      if (prog.stages.empty()) continue;
      dev::BatchJob job{};
      job.state_off = state_off[c];
      job.stage_off = (uint32_t)stages.size();
      job.rop_off = (uint32_t)rops.size();
      job.top_off = (uint32_t)tops.size();
      job.mat_off = (uint32_t)mats.size();
      job.P = prog.P;
      stages.insert(stages.end(), prog.stages.begin(), prog.stages.end());
      rops.insert(rops.end(), prog.rops.begin(), prog.rops.end());
      tops.insert(tops.end(), prog.tops.begin(), prog.tops.end());
      mats.insert(mats.end(), prog.mats.begin(), prog.mats.end());
      if (jobs.size() <= p) {
        jobs.resize(p + 1);
        job_circuit.resize(p + 1);
      }
      jobs[p].push_back(job);
      job_circuit[p].push_back(c);
    }
  }
  // ---- launches: pass p of every circuit that has one
  for (size_t p = 0; p < jobs.size(); ++p) {
    auto &js = jobs[p];
    if (js.empty()) continue;
    uint32_t blocks = 0, tmax = 0;
    std::vector<uint32_t> block_job;
    for (size_t j = 0; j < js.size(); ++j) {
      js[j].first_block = blocks;
      const uint64_t tiles = (uint64_t(1) << circ_[job_circuit[p][j]].n) >> js[j].P.t;
      if (blocks + tiles > (uint64_t)std::numeric_limits<int32_t>::max()) throw ValueError("batch: too many tiles in one launch");
      for (uint64_t b = 0; b < tiles; ++b) block_job.push_back((uint32_t)j);
      blocks += (uint32_t)tiles;
      tmax = std::max(tmax, js[j].P.t);
    }
    d_jobs_.push_back(to_device(js));
    d_block_job_.push_back(to_device(block_job));
    launch_jobs_.push_back(js);
    launch_blocks_.push_back(blocks);
    launch_tmax_.push_back(tmax);
  }
  // ---- energies: one job per circuit, chunks of up to 4096 amplitudes per workgroup
  std::vector<dev::EnergyJob> ejobs(circ_.size());
  std::vector<dev::BatchTerm> terms;
  uint32_t slots = 0;
  for (size_t c = 0; c < circ_.size(); ++c) {
    dev::EnergyJob &e = ejobs[c];
    e.state_off = state_off[c];
    e.n_amps = uint64_t(1) << circ_[c].n;
    e.chunk = std::min<uint64_t>(e.n_amps, 4096);
    e.first_block = slots;
    e.n_blocks = (uint32_t)((e.n_amps + e.chunk - 1) / e.chunk);
    slots += e.n_blocks;
    e.term_off = (uint32_t)terms.size();
    e.n_terms = (uint32_t)circ_[c].terms.size();
    for (const dev::PauliTerm &t0 : circ_[c].terms) {
      // the strings act on the circuit's final layout (relabelling moves qubits)
      dev::PauliTerm t = t0;
      t.x = t.z = 0;
      for (uint32_t q = 0; q < circ_[c].n; ++q) {
        t.x |= ((t0.x >> q) & 1u) << final_pos[c][q];
        t.z |= ((t0.z >> q) & 1u) << final_pos[c][q];
      }
      // (-1)^{popcount((i ^ x) & z)} = (-1)^{popcount(i & z)} (-1)^{popcount(x & z)}
      double s = (__builtin_popcountll(t.x & t.z) & 1) ? -1.0 : 1.0;
      // pairs (i, i ^ x): half the amplitudes, twice the weight
      const uint32_t h = t.x ? 1u << (63 - __builtin_clzll(t.x)) : 0u;
      if (t.x) s *= 2.0;
      terms.push_back(dev::BatchTerm{(uint32_t)t.x, (uint32_t)t.z, h, 0u, s * t.c_re, s * t.c_im});
    }
  }
  // Workgroup w runs on XCD w mod 8.  Every circuit goes whole to one XCD (largest first, to the
  // least loaded), so its partner reads stay in that XCD's L2; the XCDs' workgroup lists are
  // interleaved and padded with idle workgroups to equal length.
  constexpr int kXcds = 8;
  std::vector<size_t> order(circ_.size());
  for (size_t c = 0; c < order.size(); ++c) order[c] = c;
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return circ_[a].n > circ_[b].n; });
  std::vector<std::vector<dev::EnergyBlock>> per_xcd(kXcds);
  std::vector<uint64_t> load(kXcds, 0);
  for (size_t c : order) {
    const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += ejobs[c].n_amps * std::max<uint32_t>(ejobs[c].n_terms, 1);
    for (uint32_t b = 0; b < ejobs[c].n_blocks; ++b) per_xcd[x].push_back(dev::EnergyBlock{(uint32_t)c, b});
  }
  size_t longest = 0;
  for (auto &v : per_xcd) longest = std::max(longest, v.size());
  std::vector<dev::EnergyBlock> eblock;
  for (size_t k = 0; k < longest; ++k)
    for (int x = 0; x < kXcds; ++x)
      eblock.push_back(k < per_xcd[x].size() ? per_xcd[x][k] : dev::EnergyBlock{dev::kNoJob, 0});
  e_blocks_ = (uint32_t)eblock.size();
  // ---- device buffers
  const size_t ab = c64_ ? 8 : 16;
  QS_HIP(hipMalloc(&psi_, total_amps_ * ab));
  d_state_off_ = to_device(state_off);
  d_stages_ = to_device(stages);
  d_rops_ = to_device(rops);
  two_qubit_ops_ = false;
  for (const dev::RegOp &op : rops) {
    const uint32_t kind = (op.code >> 8) & 15u;
    two_qubit_ops_ = two_qubit_ops_ || kind == dev::kOp2General || kind == dev::kOp2Diag;
  }
  d_tops_ = to_device(tops);
  if (c64_) {
    std::vector<float2> m(mats.size());
    for (size_t i = 0; i < m.size(); ++i) m[i] = make_float2((float)mats[i].real(), (float)mats[i].imag());
    d_mats_ = to_device(m);
  } else {
    std::vector<double2> m(mats.size());
    for (size_t i = 0; i < m.size(); ++i) m[i] = make_double2(mats[i].real(), mats[i].imag());
    d_mats_ = to_device(m);
  }
  d_ejobs_ = to_device(ejobs);
  d_eblock_ = to_device(eblock);
  d_terms_ = to_device(terms);
  QS_HIP(hipMalloc(&d_partials_, std::max<size_t>(slots, 1) * sizeof(double)));
  QS_HIP(hipMalloc(&d_energies_, circ_.size() * sizeof(double)));
  // ---- capture one run into a graph
  QS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
  try {
    enqueue();
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(stream_, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  QS_HIP(hipStreamEndCapture(stream_, &graph_));
  QS_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
}

void Batch::enqueue() {
  auto go = [&](auto tag) {
    using R = decltype(tag);
    amp_t<R> *psi = (amp_t<R> *)psi_;
    dev::batch_init<R>(psi, total_amps_, d_state_off_, (int)circ_.size(), stream_);
    for (size_t l = 0; l < launch_jobs_.size(); ++l)
      dev::batch_tile_pass<R>(psi, d_jobs_[l], d_block_job_[l], launch_blocks_[l], launch_tmax_[l], (int)shape_.rb,
                              (const dev::TileStage *)d_stages_, (const dev::RegOp *)d_rops_,
                              (const dev::TileOp *)d_tops_, (const amp_t<R> *)d_mats_, two_qubit_ops_, stream_);
    dev::batch_energy<R>(psi, d_ejobs_, d_eblock_, e_blocks_, d_terms_, d_partials_, stream_);
  };
  if (c64_) go(float{});
  else go(double{});
  dev::batch_energy_finish(d_ejobs_, (int)circ_.size(), d_partials_, d_energies_, stream_);
}

double Batch::run(bool graph, double *energies) {
  QS_HIP(hipEventRecord(ev0_, stream_));
  if (graph) QS_HIP(hipGraphLaunch(exec_, stream_));
  else enqueue();
  QS_HIP(hipEventRecord(ev1_, stream_));
  QS_HIP(hipMemcpyAsync(h_energies_, d_energies_, circ_.size() * sizeof(double), hipMemcpyDeviceToHost, stream_));
  QS_HIP(hipStreamSynchronize(stream_));
  float ms = 0.f;
  QS_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  std::memcpy(energies, h_energies_, circ_.size() * sizeof(double));
  return ms;
}

}  // namespace qs
