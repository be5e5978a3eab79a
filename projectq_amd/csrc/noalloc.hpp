// Expectation value on a caller-owned device state with no allocation (noalloc.hip).
#pragma once

#include <cstddef>
#include <cstdint>

namespace qs {

// Pauli strings per call; the scratch array holds their records, partial sums and results.
constexpr size_t kNoAllocMaxTerms = 4096;
constexpr size_t kNoAllocScratchBytes = 32 * kNoAllocMaxTerms + 16 * 2048 + 16 * kNoAllocMaxTerms;

// sum_t coeffs[t] * Re<psi|P_t|psi> for the state of 2^n_qubits amplitudes at `state` (this GPU's
// memory; complex64 if c64, else complex128).  String t owns ops [term_offsets[t],
// term_offsets[t+1]) of (term_idx local index, term_ops 'X'|'Y'|'Z'); local index l is the qubit
// at bit position positions[l].  Synchronous on the null stream.  Throws ValueError on bad input.
double expectation_noalloc(const void *state, unsigned n_qubits, bool c64, const uint32_t *term_idx,
                           const uint8_t *term_ops, const size_t *term_offsets, size_t n_terms, const double *coeffs,
                           const uint32_t *positions, size_t n_positions);

}  // namespace qs
