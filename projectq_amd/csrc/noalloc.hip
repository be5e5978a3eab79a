// Expectation value of Pauli strings on a caller-owned device state without any allocation.
//
// The same computation as Engine::get_expectation_value (the same kernels, in the same order, with
// the same host-side combination), for callers that cannot allocate.  Its uploaded strings,
// partial sums and per-string results live in `scratch`, one file-scope device array per GPU,
// whose address on each device is looked up once (hipGetSymbolAddress) with that device current.
// Host-side work uses fixed-size stack chunks: no device or heap allocation happens in a call.
// Calls that overlap on one GPU share that GPU's array and race.
#include "noalloc.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <mutex>

#include "kernels.hpp"
#include "planner.hpp"

namespace qs {

namespace {

// scratch layout (byte offsets, 256-byte aligned)
constexpr size_t kTermsOff = 0;                                                         // PauliTerm[kNoAllocMaxTerms]
constexpr size_t kPartialsOff = kTermsOff + sizeof(dev::PauliTerm) * kNoAllocMaxTerms;  // double[2 * kReduceGrid]
constexpr size_t kResultsOff = kPartialsOff + sizeof(double) * 2 * dev::kReduceGrid;    // double[2 * kNoAllocMaxTerms]
constexpr size_t kScratchBytes = kResultsOff + sizeof(double) * 2 * kNoAllocMaxTerms;
static_assert(sizeof(dev::PauliTerm) == 32, "PauliTerm record size");
static_assert(kPartialsOff % 256 == 0 && kResultsOff % 256 == 0, "scratch regions stay 256-byte aligned");
static_assert(kScratchBytes == kNoAllocScratchBytes, "scratch size as declared in noalloc.hpp");

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
char *g_scratch[kMaxDevices] = {};

// This GPU's copy of `scratch`, looked up on the first call with the device current.  The name
// `scratch` is a host-side stand-in: only the looked-up address may reach kernels and copies.
char *scratch_on_current_device() {
  int dev = 0;
  QS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw DeviceError("expectation_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    QS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw DeviceError("expectation_noalloc: no device address for the scratch array");
    g_scratch[dev] = static_cast<char *>(p);
  }
  return g_scratch[dev];
}

constexpr size_t kChunk = 128;  // strings encoded per host chunk (stack storage)

// String t as  omega * X^x Z^z  (the engine's encoding), with its local indices mapped.
void encode(size_t t, const uint32_t *term_idx, const uint8_t *term_ops, const size_t *term_offsets,
            const uint32_t *positions, size_t n_positions, dev::PauliTerm *pt, std::complex<double> *omega) {
  PauliEncoder e;
  for (size_t j = term_offsets[t]; j < term_offsets[t + 1]; ++j) {
    if (term_idx[j] >= n_positions) throw ValueError("qubit_operator acts on more qubits than contained in the qureg.");
    const char op = (char)term_ops[j];
    if (op != 'X' && op != 'Y' && op != 'Z') throw ValueError("Pauli operator must be X, Y or Z");
    e.add(positions[term_idx[j]], op);
  }
  pt->x = e.x;
  pt->z = e.z;
  *omega = e.omega;
}

}  // namespace

double expectation_noalloc(const void *state, unsigned n_qubits, bool c64, const uint32_t *term_idx,
                           const uint8_t *term_ops, const size_t *term_offsets, size_t n_terms, const double *coeffs,
                           const uint32_t *positions, size_t n_positions) {
  if (n_terms > kNoAllocMaxTerms)
    throw ValueError("expectation_noalloc: at most 4096 Pauli strings per call (the scratch capacity)");
  if (n_qubits > 40) throw ValueError("expectation_noalloc: at most 40 qubits");
  if (!state) throw ValueError("expectation_noalloc: the state pointer is null");
  for (size_t l = 0; l < n_positions; ++l)
    if (positions[l] >= n_qubits) throw ValueError("expectation_noalloc: a qubit position lies outside the state");
  // validate every string before any device work
  for (size_t t = 0; t < n_terms; ++t) {
    dev::PauliTerm pt;
    std::complex<double> w;
    encode(t, term_idx, term_ops, term_offsets, positions, n_positions, &pt, &w);
  }
  if (n_terms == 0) return 0.0;
  const uint64_t n_amps = uint64_t(1) << n_qubits;
  char *base = scratch_on_current_device();
  dev::PauliTerm *d_terms = reinterpret_cast<dev::PauliTerm *>(base + kTermsOff);
  double *d_partials = reinterpret_cast<double *>(base + kPartialsOff);
  double *d_results = reinterpret_cast<double *>(base + kResultsOff);
  const hipStream_t s = nullptr;  // the null stream; the call returns when its work is done
  QS_HIP(hipStreamSynchronize(s));
  dev::PauliTerm pt[kChunk];
  std::complex<double> omega[kChunk];
  // the engine's choice: one pass for all strings while the state fits the last-level cache, one
  // streaming pass per string above it
  const bool cached = n_amps * (c64 ? 8u : 16u) <= (uint64_t(64) << 20);
  if (cached) {
    for (size_t t0 = 0; t0 < n_terms; t0 += kChunk) {
      const size_t m = std::min(kChunk, n_terms - t0);
      for (size_t i = 0; i < m; ++i) {
        encode(t0 + i, term_idx, term_ops, term_offsets, positions, n_positions, &pt[i], &omega[i]);
        if (!dev::pauli_hermitian(pt[i].x, pt[i].z, omega[i])) {
          pt[i] = dev::PauliTerm{0, 0, 0.0, 0.0};  // Re<psi|P|psi> = 0 for an anti-Hermitian string
          continue;
        }
        const std::complex<double> c = coeffs[t0 + i] * omega[i];
        pt[i].c_re = c.real();
        pt[i].c_im = c.imag();
      }
      QS_HIP(hipMemcpy(d_terms + t0, pt, m * sizeof(dev::PauliTerm), hipMemcpyHostToDevice));
    }
    if (c64)
      dev::pauli_energy<float>(static_cast<const float2 *>(state), n_amps, d_terms, (int)n_terms, d_partials, s);
    else
      dev::pauli_energy<double>(static_cast<const double2 *>(state), n_amps, d_terms, (int)n_terms, d_partials, s);
    dev::finish_sum(d_partials, dev::kReduceGrid, 1, d_results, s);
    double out = 0.0;
    QS_HIP(hipMemcpy(&out, d_results, sizeof(double), hipMemcpyDeviceToHost));
    return out;
  }
  const BitInsert none{};
  double total = 0.0;
  for (size_t t0 = 0; t0 < n_terms; t0 += kChunk) {
    const size_t m = std::min(kChunk, n_terms - t0);
    for (size_t i = 0; i < m; ++i) {
      encode(t0 + i, term_idx, term_ops, term_offsets, positions, n_positions, &pt[i], &omega[i]);
      double *out = d_results + 2 * (t0 + i);
      if (!dev::pauli_hermitian(pt[i].x, pt[i].z, omega[i])) continue;  // contributes 0, read back unused
      if (pt[i].x == 0 && pt[i].z == 0) {
        if (c64)
          dev::norm_masked<float>(static_cast<const float2 *>(state), n_amps, none, 0, d_partials, s);
        else
          dev::norm_masked<double>(static_cast<const double2 *>(state), n_amps, none, 0, d_partials, s);
        dev::finish_sum(d_partials, dev::kReduceGrid, 1, out, s);
      } else {
        if (c64)
          dev::pauli_expect<float>(static_cast<const float2 *>(state), n_amps, pt[i].x, pt[i].z, d_partials, s);
        else
          dev::pauli_expect<double>(static_cast<const double2 *>(state), n_amps, pt[i].x, pt[i].z, d_partials, s);
        dev::finish_sum(d_partials, dev::kReduceGrid, 2, out, s);
      }
    }
    double res[2 * kChunk];
    QS_HIP(hipMemcpy(res, d_results + 2 * t0, 2 * m * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < m; ++i) {
      double delta;
      if (!dev::pauli_hermitian(pt[i].x, pt[i].z, omega[i]))
        delta = 0.0;
      else if (pt[i].x == 0 && pt[i].z == 0)
        delta = omega[i].real() * res[2 * i];
      else
        delta = (omega[i] * std::complex<double>(res[2 * i], res[2 * i + 1])).real();
      total += coeffs[t0 + i] * delta;
    }
  }
  return total;
}

}  // namespace qs
