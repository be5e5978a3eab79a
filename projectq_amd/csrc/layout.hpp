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
};

// One op of a register stage.  code = kind << 8 | b1 << 4 | b0 with b0 (< b1) register bits:
//   kind 0: 1-qubit general 2x2 on b0      kind 3: 1-qubit X (swap) on b0
//   kind 1: 1-qubit diagonal on b0         kind 4: 2-qubit general 4x4 on (b0, b1)
//   kind 2: 1-qubit anti-diagonal on b0    kind 5: 2-qubit diagonal on (b0, b1)
//   kind 6: 1-qubit real 2x2 on b0
// Matrix bit 0 belongs to b0 and bit 1 to b1.
enum : uint32_t {
  kOp1General = 0,
  kOp1Diag = 1,
  kOp1Anti = 2,
  kOp1X = 3,
  kOp2General = 4,
  kOp2Diag = 5,
  kOp1Real = 6
};
struct RegOp {
  uint32_t code;
  uint32_t rcmask;  // controls among the stage's register bits (register-index mask)
  uint32_t lcmask;  // tile-local controls outside the stage's register bits
  uint32_t mat;     // first element of the op's matrix in the pass's matrix array
  uint64_t gcmask;  // controls outside the tile
};

struct TilePass {
  uint32_t t, L, h, n_ops;
  uint32_t n_stages;
  uint32_t hpos[kTileMaxHigh];  // global positions of the high tile bits, ascending
  BitInsert tiles;              // all tile positions (global, ascending): enumerates tile bases
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

// Expectation value of one circuit inside a batched launch (k_batch_energy): workgroups
// first_block .. first_block + n_blocks - 1 each cover `chunk` amplitudes.
struct EnergyJob {
  uint64_t state_off;
  uint64_t n_amps;
  uint64_t chunk;
  uint32_t first_block, n_blocks;
  uint32_t term_off, n_terms;
};

// LDS address of tile-local index x: an XOR swizzle of the low bits by the next bits, so that
// threads reading with power-of-two strides spread over the banks.  Pairs (2j, 2j+1) stay
// adjacent for complex64 so 16-byte vector accesses still work.  A bijection on [0, 2^t).
inline constexpr uint32_t lds_slot_c64(uint32_t x) { return x ^ (((x >> 5) & 15u) << 1); }
inline constexpr uint32_t lds_slot_c128(uint32_t x) { return x ^ ((x >> 4) & 15u); }

}  // namespace dev
}  // namespace qs
