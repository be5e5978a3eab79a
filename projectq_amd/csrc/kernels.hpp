// Launchers of the HIP kernels (kernels.hip).  All take the simulator's stream; none
// synchronises.  Index conventions: amplitude index bit p belongs to the qubit at position p.
#pragma once

#include <complex>

#include "common.hpp"

namespace qs {
namespace dev {

constexpr int kBlock = 256;       // threads per workgroup (4 wave64)
constexpr int kReduceGrid = 2048; // workgroups of every reduction -> fixed partial-sum count

// One Pauli string as  omega * X^x Z^z  with omega folded into the coefficient.
struct PauliTerm {
  uint64_t x, z;
  double c_re, c_im;
};

// omega X^x Z^z (omega one of 1, -1, i, -i) is Hermitian iff omega^2 = (-1)^popcount(x & z); the
// product of Paulis on distinct qubits always is, a string naming a qubit twice may be
// anti-Hermitian, and then Re<psi|P|psi> = 0.
inline bool pauli_hermitian(uint64_t x, uint64_t z, std::complex<double> omega) {
  const double sigma = (__builtin_popcountll(x & z) & 1) ? -1.0 : 1.0;
  return (omega * omega).real() * sigma > 0.0;
}

// Register layout and constants of one emulate_math call.
constexpr int kMaxRegs = 8;
struct MathArgs {
  int32_t opcode;
  int32_t n_regs;
  int64_t a, N;
  uint64_t cmask;
  uint8_t width[kMaxRegs];
  uint8_t pos[kMaxRegs][32];
};

// tile passes (TilePass / TileStage / RegOp / TileOp in layout.hpp); rb = register bits (4, 5);
// two_qubit_ops: whether any register op of the pass is a 2-qubit one (false lets complex64
// passes run a kernel without the 2-qubit op forms)
template <typename R>
void tile_pass(amp_t<R> *psi, uint64_t n_tiles, const TilePass &P, int rb, const TileStage *d_stages,
               const RegOp *d_rops, const TileOp *d_tops, const amp_t<R> *d_mats, bool two_qubit_ops, hipStream_t s);

// set the dynamic-LDS limit of every tile-pass kernel (call before any graph capture)
void prepare_tile_kernels();

// ---- batched circuits (one state per circuit in one buffer) ----
// every state <- |0...0>: zero the whole buffer, then amplitude state_off[c] = 1 (no memset, so
// the sequence can be captured in a graph)
template <typename R>
void batch_init(amp_t<R> *psi, uint64_t total_amps, const uint64_t *state_off, int n_states, hipStream_t s);
// one pass of every circuit: n_blocks workgroups, block_job maps each to its job;
// two_qubit_ops: whether any register op of the batch is a 2-qubit one
template <typename R>
void batch_tile_pass(amp_t<R> *psi, const BatchJob *jobs, const uint32_t *block_job, uint32_t n_blocks,
                     uint32_t max_t, int rb, const TileStage *stages, const RegOp *rops, const TileOp *tops,
                     const amp_t<R> *mats, bool two_qubit_ops, hipStream_t s);
// partials[b] for every energy workgroup, then energies[j] = sum of job j's partials in order
template <typename R>
void batch_energy(const amp_t<R> *psi, const EnergyJob *jobs, const EnergyBlock *blocks, uint32_t n_blocks,
                  const BatchTerm *terms, double *partials, hipStream_t s);
void batch_energy_finish(const EnergyJob *jobs, int n_jobs, const double *partials, double *energies,
                         hipStream_t s);

// dense gate block: psi[base + off(r)] = sum_c M[r][c] psi[base + off(c)] for every base with
// target bits 0 and control bits 1.  tpos[l] is the position of matrix bit l; ins lists the
// sorted target+control positions; n_groups = 2^(n - k - #controls).
template <typename R>
void apply_block(amp_t<R> *psi, uint64_t n_groups, const BitInsert &ins, uint64_t cmask,
                 const uint32_t *tpos, int k, const amp_t<R> *d_mat, hipStream_t s);

// the same dense block for a complex64 state on the matrix cores (k = 3, 4, 5): d_table is the
// per-lane A-operand table of mfma::operand_table (mfma_layout.hpp)
void apply_block_mfma(float2 *psi, uint64_t n_groups, const BitInsert &ins, uint64_t cmask, const uint32_t *tpos,
                      int k, const float *d_table, hipStream_t s);

// partials[0..kReduceGrid) <- block sums of |psi_i|^2 over i = insert(g) | val, g < n_sub
template <typename R>
void norm_masked(const amp_t<R> *psi, uint64_t n_sub, const BitInsert &ins, uint64_t val,
                 double *partials, hipStream_t s);
// out[0..width) <- sum over blocks of partials[.. * width + j], fixed order (deterministic)
void finish_sum(const double *partials, int n_partials, int width, double *out, hipStream_t s);

// The three measurement kernels cover a subspace of the state: compact index c < n_sub stands for
// amplitude insert_zeros(c, sub) | sval (the bits in `sub` fixed to their values in sval; the
// caller knows every amplitude outside the subspace is zero).  sub.count = 0, sval = 0: the state.

// psi_i <- 0 where (i & mask) != val, psi_i * scale elsewhere, for i in the subspace
template <typename R>
void collapse(amp_t<R> *psi, uint64_t n_sub, const BitInsert &sub, uint64_t sval, uint64_t mask, uint64_t val,
              double scale, hipStream_t s);

// sums[c] <- sum |psi_i|^2 over chunk c of 2^chunk_log2 consecutive compact indices
template <typename R>
void chunk_norms(const amp_t<R> *psi, uint64_t n_sub, const BitInsert &sub, uint64_t sval, int chunk_log2,
                 double *sums, hipStream_t s);
// *pick <- first compact index j in [start, start+len) with base + sum_{start<=c<=j} |psi_c|^2 >=
// rnd (left at its initial value if none)
template <typename R>
void chunk_pick(const amp_t<R> *psi, const BitInsert &sub, uint64_t sval, uint64_t start, uint64_t len, double base,
                double rnd, unsigned long long *pick, hipStream_t s);

// flags[0] |= any |psi_i|^2 > tol with bit p clear; flags[1] likewise with bit p set
template <typename R>
void classical_flags(const amp_t<R> *psi, uint64_t n_amps, int p, double tol, int *flags,
                     hipStream_t s);
// *key <- min over pairs q (index with bit p removed) holding an entry > tol of
//         2q + (entry with bit p clear is <= tol)
template <typename R>
void first_pair(const amp_t<R> *psi, uint64_t n_pairs, int p, double tol,
                unsigned long long *key, hipStream_t s);
// dst[j] <- src[j with a bit `value` inserted at p], j < n_out
template <typename R>
void drop_bit(const amp_t<R> *src, amp_t<R> *dst, uint64_t n_out, int p, int value,
              hipStream_t s);
// In place: exchange index bits a and b (a != b) of every amplitude of a 2^n state
template <typename R>
void swap_bits(amp_t<R> *psi, uint64_t n_amps, int a, int b, hipStream_t s);
// dst <- src with its index bits permuted, one pass through LDS tiles (PermTile)
template <typename R>
void permute_tiles(const amp_t<R> *src, amp_t<R> *dst, uint64_t n_amps, const PermTile &P, hipStream_t s);

// Both expectation kernels take Hermitian strings only, omega X^x Z^z with omega^2 =
// (-1)^popcount(x & z) (pauli_hermitian); the caller drops the others (their real part is zero).
// partials[2*b + {0,1}] <- block sums whose total S satisfies Re(omega S) = Re<psi|omega X^x Z^z|psi>
template <typename R>
void pauli_expect(const amp_t<R> *psi, uint64_t n_amps, uint64_t x, uint64_t z,
                  double *partials, hipStream_t s);
// partials[b] <- block sums of sum_t Re<psi|c_t X^x_t Z^z_t|psi> (all terms in one pass; c_t
// already carries the string's phase omega_t)
template <typename R>
void pauli_energy(const amp_t<R> *psi, uint64_t n_amps, const PauliTerm *terms, int n_terms,
                  double *partials, hipStream_t s);
// dst_i <- f * sum_t c_t (-1)^popcount((i^x_t)&z_t) src_{i^x_t}; with `evolve` also
// out_i += dst_i and nrm partial sums of |dst_i|^2 on entries with (i & cmask) == cmask
template <typename R>
void pauli_sum(const amp_t<R> *src, amp_t<R> *dst, amp_t<R> *out, uint64_t n_amps,
               const PauliTerm *terms, int n_terms, double f_re, double f_im, bool evolve,
               uint64_t cmask, double *partials, hipStream_t s);
// psi_i *= c where (i & cmask) == cmask
template <typename R>
void scale_masked(amp_t<R> *psi, uint64_t n_amps, uint64_t cmask, double c_re, double c_im,
                  hipStream_t s);

// emulate_math: dest(i) for every i.  dst[dest(i)] = src[i] (caller guarantees injectivity);
// err |= 1 when a result leaves int32.  With `table` the register outputs come from
// table[packed inputs * n_regs + r] instead of the opcode.
template <typename R>
void math_permute(const amp_t<R> *src, amp_t<R> *dst, uint64_t n_amps, const MathArgs &args,
                  const int32_t *table, int *err, hipStream_t s);
// bitmap (n_amps bits, zeroed by caller) marks dest(i); collide |= 1 on a repeated target
void math_injective(uint64_t n_amps, const MathArgs &args, const int32_t *table,
                    uint32_t *bitmap, int *collide, int *err, hipStream_t s);
// general (non-injective) form: dst[t] = sum over i with dest(i) = t, in increasing i
// (stable sort of (dest, i) pairs, then ordered segment sums).  dst must be zeroed.
template <typename R>
void math_accumulate(const amp_t<R> *src, amp_t<R> *dst, uint64_t n_amps, const MathArgs &args,
                     const int32_t *table, int *err, hipStream_t s);

// partials[2*b + {0,1}] <- block sums of conj(a_i) * b_i (either operand c128 or c64)
void overlap(const void *a, bool a_c64, const void *b, bool b_c64, uint64_t n_amps,
             double *partials, hipStream_t s);

// partials[2*b] <- block max of |a_i - b_i|, partials[2*b + 1] <- block count of i with
// a_i != b_i (stored values compared exactly); either operand c128 or c64
void compare(const void *a, bool a_c64, const void *b, bool b_c64, uint64_t n_amps, double *partials,
             hipStream_t s);

// dst[i] = src[i] converted between complex128 and complex64 (round to nearest on narrowing)
template <typename S, typename D>
void convert_amps(const amp_t<S> *src, amp_t<D> *dst, uint64_t n, hipStream_t s);

}  // namespace dev
}  // namespace qs
