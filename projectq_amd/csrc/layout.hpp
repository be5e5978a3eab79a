// Plain-data layouts shared by host planning code and the HIP kernels (no HIP dependency, so
// host-only tools can include it).
#pragma once

#include <cstdint>

namespace qs {

// Largest number of bit positions a kernel argument can name (targets + controls + ...).
constexpr int kMaxBits = 48;

// Sorted list of bit positions at which a zero is inserted into a compact counter; used to
// enumerate every index whose chosen bits are fixed (targets 0, controls 1, ...).
struct BitInsert {
  uint32_t count;
  uint8_t pos[kMaxBits];
};

// One pass of an index-bit permutation through LDS tiles.  A tile holds the 2^m amplitudes whose
// source bits outside `src_pos` are fixed; it is loaded with tile bit r at source position
// src_pos[r] (the first bits are the source's lowest positions: contiguous reads) and stored with
// tile bit r at destination position dst_pos[r] (the first bits are the destination's lowest
// positions: contiguous writes), store tile bit r taking load tile bit st_from[r].  The tile id's
// bits go to the source positions outer_src[i] and the destination positions outer_dst[i].
constexpr int kPermTileMaxBits = 12;
struct PermTile {
  uint32_t m, n_outer;
  uint8_t src_pos[kPermTileMaxBits], dst_pos[kPermTileMaxBits], st_from[kPermTileMaxBits];
  uint8_t outer_src[kMaxBits], outer_dst[kMaxBits];
};


namespace dev {

// ---- tile passes: one HBM read + write of the state, many gates applied in LDS ----------------
// A tile is every amplitude whose non-tile bits are fixed: 2^t of them, stored in LDS at the
// tile-local index (low L global bits, then the h high tile positions in ascending order).
constexpr int kTileMaxHigh = 9;
constexpr int kTileMaxBits = 13;
struct TileOp {
  uint32_t k;          // targets (1..5)
  uint32_t nfixed;     // targets + tile-local controls
  uint32_t lcmask;     // tile-local control bits (must be 1)
  uint32_t mat;        // first element of this op's 2^k x 2^k matrix in the pass's matrix array
  uint64_t gcmask;     // control bits outside the tile (uniform per tile)
  uint32_t fixed[16];  // tile-local positions of targets + local controls, ascending
  uint32_t off[32];    // tile-local offset of matrix index j
};

// In-tile execution is split into stages.  A register stage holds 2^RB amplitudes per thread
// (the tile-local bits `fixed[0..RB)` vary inside a thread) and applies a run of 1- and
// 2-qubit ops to them without touching LDS; an LDS stage applies one TileOp of any width
// directly on the LDS tile.  RB = 5 for complex64, 4 for complex128.
enum : uint32_t { kStageRegisters = 0, kStageLds = 1 };
struct TileStage {
  uint32_t kind;        // kStageRegisters or kStageLds
  uint32_t begin, end;  // RegOp index range (register stage) or the TileOp index (LDS stage)
  uint32_t fixed[5];    // register stage: tile-local position of register bit i, ascending
  uint32_t soff[32];    // register stage: LDS byte offset slot(index of register i), i < 2^RB
};

// One op of a register stage.  code = cb << 12 | kind << 8 | b1 << 4 | b0 with b0 (< b1)
// register bits and cb = 1 + the op's one control among the register bits (0: none; an op with
// two or more register-bit controls runs as an LDS op instead):
//   kind 0: 1-qubit general 2x2 on b0      kind 3: 1-qubit X (swap) on b0
//   kind 1: 1-qubit diagonal on b0         kind 4: 2-qubit general 4x4 on (b0, b1)
//   kind 2: 1-qubit anti-diagonal on b0    kind 5: 2-qubit diagonal on (b0, b1)
//   kind 6: 1-qubit real 2x2 on b0         kind 7: 1-qubit Z (diag(1, -1): sign flip) on b0
// Matrix bit 0 belongs to b0 and bit 1 to b1.
enum : uint32_t {
  kOp1General = 0,
  kOp1Diag = 1,
  kOp1Anti = 2,
  kOp1X = 3,
  kOp2General = 4,
  kOp2Diag = 5,
  kOp1Real = 6,
  kOp1Z = 7
};
struct RegOp {
  uint32_t code;
  uint32_t rcmask;  // controls among the stage's register bits (register-index mask, <= 1 bit)
  uint32_t lcmask;  // tile-local controls outside the stage's register bits
  uint32_t mat;     // first element of the op's matrix in the pass's matrix array
  uint64_t gcmask;  // controls outside the tile
};

struct TilePass {
  uint32_t t, L, h, n_ops;
  uint32_t n_stages;
  uint32_t hpos[kTileMaxHigh];  // global positions of the high tile bits, ascending
  BitInsert tiles;              // all tile positions (global, ascending): enumerates tile bases
  // Qubit relabelling at the store: tile-local bit b of the stored tile takes bit sperm[b] of the
  // tile as computed (sperm[b] = b: no relabelling).  A product of bit exchanges, so the stored
  // amplitude at tile-local index l is the computed one at sum_b bit_b(l) << sperm[b].
  uint32_t sperm[kTileMaxBits];
};

// One pass of one circuit inside a batched launch (k_batch_tile_pass).  Offsets index the
// batch-wide arrays of stages, register ops, LDS ops and matrices.
struct BatchJob {
  uint64_t state_off;    // first amplitude of this circuit's state
  uint32_t first_block;  // first workgroup of this job in the launch
  uint32_t stage_off, rop_off, top_off, mat_off;
  uint32_t pad;
  TilePass P;
};

// Expectation value of one circuit inside a batched launch (k_batch_energy): its amplitudes in
// n_blocks chunks of `chunk`, whose partial sums go to partials[first_block + chunk index].
struct EnergyJob {
  uint64_t state_off;
  uint64_t n_amps;
  uint64_t chunk;
  uint32_t first_block, n_blocks;
  uint32_t term_off, n_terms;
};

// One workgroup of the batched energy launch: chunk `chunk` of energy job `job`; kNoJob marks an
// idle workgroup (the launch pads every XCD's share of workgroups to the same length).
constexpr uint32_t kNoJob = 0xffffffffu;
struct EnergyBlock {
  uint32_t job, chunk;
};

// A Pauli string of a batched circuit (at most 30 qubits) with the constant part of its sign
// folded into the coefficient: (P psi)_i = c (-1)^{popcount(i & z)} psi_{i ^ x}.  For x != 0 the
// amplitudes i and i ^ x contribute complex-conjugate terms to <psi|P|psi> (P is Hermitian, the
// coefficient real), so only the i with bit `h` (the highest bit of x) clear are summed and the
// coefficient carries the factor 2; for x = 0, h = 0 and every i is summed.
struct BatchTerm {
  uint32_t x, z, h, pad;
  double c_re, c_im;
};

// LDS slot of tile-local index x: an XOR swizzle that spreads the accesses of a register stage
// over the LDS banks.  In a register stage the lanes of a wave vary over the lowest tile bits
// that are not register bits, so the bank of a lane is a GF(2)-linear function of those bits;
// the swizzle adds to the low (bank) bits a mask for every set bit b >= 4 (c64) / >= 3 (c128).
// gfx950 LDS has 64 banks of 4 bytes; it serves 32 lanes of 8-byte accesses (bank pair = 8-byte
// unit mod 32) or 16 lanes of 16-byte accesses (bank quad = 16-byte unit mod 16) per cycle.
// The masks were picked by a search over all 495 four-bit register sets of a 12-bit tile,
// minimising the mean conflict degree of those accesses (c64: 8-byte amplitudes, c128: 16-byte).  Bit b only flips bits below min(b, 5), so
// the map is unitriangular: linear, a bijection on [0, 2^t), and slot(a | b) = slot(a) ^ slot(b)
// for indices without common bits.  complex64 pairs (2j, 2j+1) stay adjacent but may swap.
inline constexpr uint32_t lds_slot_c64(uint32_t x) {
  return x ^ (((x >> 4) & 1u) * 11u) ^ (((x >> 5) & 1u) * 30u) ^ (((x >> 6) & 1u) * 15u) ^
         (((x >> 7) & 1u) * 23u) ^ (((x >> 8) & 1u) * 16u) ^ (((x >> 9) & 1u) * 3u) ^
         (((x >> 10) & 1u) * 31u) ^ (((x >> 11) & 1u) * 23u);
}
inline constexpr uint32_t lds_slot_c128(uint32_t x) {
  return x ^ (((x >> 3) & 1u) * 6u) ^ (((x >> 4) & 1u) * 13u) ^ (((x >> 5) & 1u) * 7u) ^
         (((x >> 6) & 1u) * 11u) ^ (((x >> 7) & 1u) * 8u) ^ (((x >> 8) & 1u) * 12u) ^
         (((x >> 10) & 1u) * 11u) ^ (((x >> 11) & 1u) * 4u);
}

// Bank-conflict degree of a register stage under that model: the lowest `lanes_bits`
// non-register tile bits vary across the lanes served together (5 for complex64, 4 for
// complex128), and the bank of a lane is the low `lanes_bits` bits of its slot.
// Returns 2^(lanes_bits - rank).
inline uint32_t stage_conflict_degree(uint32_t reg_bits_mask, uint32_t t, bool c64, uint32_t lanes_bits) {
  uint32_t rows[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t n = 0;
  const uint32_t bank_mask = (1u << lanes_bits) - 1;
  for (uint32_t b = 0; b < t && n < lanes_bits; ++b)
    if (!((reg_bits_mask >> b) & 1)) rows[n++] = (c64 ? lds_slot_c64(1u << b) : lds_slot_c128(1u << b)) & bank_mask;
  uint32_t rank = 0;
  for (uint32_t bit = 0; bit < lanes_bits; ++bit) {
    uint32_t piv = n;
    for (uint32_t i = rank; i < n; ++i)
      if ((rows[i] >> bit) & 1) {
        piv = i;
        break;
      }
    if (piv == n) continue;
    const uint32_t tmp = rows[rank];
    rows[rank] = rows[piv];
    rows[piv] = tmp;
    for (uint32_t i = 0; i < n; ++i)
      if (i != rank && ((rows[i] >> bit) & 1)) rows[i] ^= rows[rank];
    ++rank;
  }
  return 1u << (lanes_bits - rank);
}

// Fill the register offsets of a register stage whose `rb` register bits are in st.fixed: the
// LDS byte offset of the tile index made of register bits i (i < 2^rb), so that the kernel
// addresses register i of a thread as slot_bytes(lb) ^ soff[i].
inline void fill_stage_offsets(TileStage &st, uint32_t rb, bool c64) {
  for (uint32_t i = 0; i < 32; ++i) {
    uint32_t o = 0;
    for (uint32_t b = 0; b < rb; ++b)
      if ((i >> b) & 1) o |= 1u << st.fixed[b];
    st.soff[i] = i < (1u << rb) ? (c64 ? lds_slot_c64(o) * 8u : lds_slot_c128(o) * 16u) : 0u;
  }
}

}  // namespace dev
}  // namespace qs
