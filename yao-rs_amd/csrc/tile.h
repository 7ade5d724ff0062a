// tile.h — multi-gate tile passes: data shared by the host planner and the kernel.
#pragma once

#include "common.h"

namespace yh {

// A tile is 2^kTileBits amplitudes: tile qubits 0..kTileLow-1 are memory bits
// 0..kTileLow-1 (a 512 B contiguous run for complex128), tile qubits
// kTileLow..kTileBits-1 are high memory bits chosen per pass.
constexpr int kTileBits = 12;
constexpr int kTileN = 1 << kTileBits;
constexpr int kTileLow = 5;
constexpr int kTileHigh = kTileBits - kTileLow;
constexpr int kTileThreads = 256;
constexpr int kGroupBits = 4;  // each thread holds a 2^4-amplitude sub-cube of a group
constexpr int kAmpsPerThread = kTileN / kTileThreads;

// One gate inside a group, in the group's local index space (local qubit k <-> tile
// qubit Group::q[k]). Gate-index bit 0 <-> t0, bit 1 <-> t1 (yao order).
enum MicroKind : uint8_t { MG_DENSE1 = 0, MG_DIAG1 = 1, MG_DENSE2 = 2, MG_PERM1 = 3 /* X: pair swap */ };
struct MicroGate {
  uint8_t kind, t0, t1, cmask, cval, pad[3];
  double m[16][2];  // dense1: 4 entries, diag1: 2, dense2: 16 (re, im)
};

struct TileGroup {
  uint8_t q[kGroupBits];  // tile qubits of the 4 local qubits (distinct)
  int32_t first, count;   // micro gates [first, first + count)
};

struct TilePass {
  uint8_t hbits[kTileHigh];  // memory bit of tile qubit kTileLow + k
  int32_t first_group, ngroups;
  int32_t mg_end;            // one past the pass's last micro gate
};

// A workgroup holds a window of this many consecutive micro gates in LDS (and the
// pass's group descriptors, up to kTileGroupsLds); the planner splits a group with
// more micro gates than the window holds.
constexpr int kTileMicroWin = 30;
constexpr int kTileGroupsLds = 64;

}  // namespace yh
