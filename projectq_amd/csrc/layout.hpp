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
struct TilePass {
  uint32_t t, L, h, n_ops;
  uint32_t hpos[kTileMaxHigh];  // global positions of the high tile bits, ascending
  BitInsert tiles;              // all tile positions (global, ascending): enumerates tile bases
};

}  // namespace dev
}  // namespace qs
