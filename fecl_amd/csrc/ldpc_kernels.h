// Device-side building blocks of the flooding BP decoder (gfx950, wave64).
//
// Layout in HBM.  Frames are processed in tiles of TW = 64*FPL frames (FPL = frames per lane,
// 1 for FP32 messages, 2 for FP16/BF16) so that every lane moves one 32-bit word per edge and
// every wavefront access is 256 contiguous bytes, whatever the Tanner graph looks like:
//   msg   [tile][edge][64] 32-bit words   edge messages, in place (v2c before a check update,
//                                         c2v after it), edges in H's row-major order
//   L     [tile][var][TW]  float          channel LLR (+ a-priori)
//   hard  [tile][var][FPL] uint64         hard decisions; bit `lane` of word k is frame
//                                         lane*FPL + k of the tile
//   phase / iter [tile*TW]                per-frame decoder state
// Graph indices (row_ptr, col, var_ptr, var_edges) are shared by all frames of a tile and are
// read with scalar loads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fecl {

constexpr int kWave = 64;
constexpr int kBlock = 256;   // 4 wavefronts per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;

// Per-frame phase (low two bits) and flags.
enum : uint8_t {
  PH_FREE = 0, PH_RUN = 1, PH_FINAL = 2, PH_DONE = 3, PH_MASK = 3,
  PH_FRESH = 4,    // v2c not yet materialised: the next check update reads L[col(e)]
  PH_VNINIT = 8,   // state supplied: a variable-node pass must build v2c from it first
};
// Tile flags: any frame RUN or FINAL (check and variable kernels work), any frame RUN
// (syndrome works), any frame waiting for its state-seeded variable pass, and "uniform": every
// frame of the tile is RUN with no pending flag and the same iteration count (tile_iter), so
// kernels need not read per-frame state.  The control kernel and the retile after a compaction
// set TF_UNIFORM from the phases; every other writer of a tile's flags clears it.
// Bits TF_GROUP_SHIFT.. hold the tile's lane groups: bit g is set when lanes
// kLaneGroup*g .. kLaneGroup*g+kLaneGroup-1 may hold a frame that is RUN or FINAL or waits for
// its state-seeded pass.  A set bit may be stale (a superset is safe); a clear bit is exact.
// Check and variable kernels skip the row loads of the other lanes (lane_for_loads).
enum : uint32_t { TF_LIVE = 1, TF_RUN = 2, TF_VNINIT = 4, TF_UNIFORM = 8 };
constexpr int TF_GROUP_SHIFT = 16, kLaneGroup = 4;

struct DevCode {
  const int32_t* row_ptr;
  const int32_t* col;
  const int32_t* var_ptr;
  const int32_t* var_edges;
  // var_tab[v * vstride + j]: j-th edge of variable v in ascending edge order, -1 past its
  // degree; vstride is the variable kernel's degree bound (4, 8, 16 or 32), so one scalar
  // load fetches a variable's whole edge list
  const int32_t* var_tab;
  int32_t vstride;
  int32_t M, N, K, E;
};

struct DevState {
  uint32_t* msg;
  float* L;
  uint64_t* hard;
  uint8_t* phase;
  int32_t* iter;
  uint64_t* fail;          // [tile][FPL]
  uint32_t* tile_flags;    // [tile]
  int32_t* tile_iter;      // [tile] iteration count of a TF_UNIFORM tile
  int32_t* active;         // [iterations] frames still iterating after each step
  int32_t tiles;
  // Eager decodes: [tile] frames still iterating in each tile after this step, written by the
  // control kernel into mapped page-locked memory that the host reads (null: not written).
  int32_t* live_out = nullptr;
};

struct DevOpts {
  const double* sf;        // [sf_rows][sf_cols]
  const float* sf32;       // same table rounded to float
  int32_t sf_rows, sf_cols;
  int32_t sf_mode;         // 0: all factors 1.0; 1: all exactly representable in float; 2: general
  int32_t iterations;
  int32_t algorithm;
  int32_t fused_syn;       // syndrome folded into the check kernel (scaling rows == 1)
  int32_t nt_rows;         // non-temporal message rows: bit 0 check kernel, bit 1 variable kernel
  float lin_y[8];          // Linear table knots y[i] = log(1 + exp(-i/2)) (float), entry 7 = 0
  float lin_s[8];          // slopes y[i+1] - y[i] (float), entry 7 = 0
};

// A compaction decided on the device (captured graphs): whether it moves frames, and how many
// frames are live.  Kernels handed one do nothing unless `go` is set.
struct CompactCtl {
  int32_t go;
  int32_t live;
};

// Outputs of a finished batch, in tile layout (device buffers owned by the engine).
struct DevOut {
  uint64_t* hard;          // [tile][K][FPL]  decode hard decisions (L + sum ascending >= 0)
  float* ext;              // [tile][N][TW]   extrinsic = sum of c2v ascending (nullable)
  float* app;              // [tile][K][TW]   L + ext (nullable)
};

}  // namespace fecl
