// Persistent multi-code decode: the work queue a single kernel launch walks (shared by the
// kernel in ldpc_persist.hip and the schedule check in tests/cpp/persist_schedule_test.cpp).
//
// One launch decodes the batches of several codes (min-sum).  Every workgroup repeatedly takes
// the next item of one global queue.  The queue lists, step after step, all check-node items
// of every code and tile, then all variable-node items, then all syndrome items (the
// workgroup that finishes a tile's last syndrome item runs that tile's per-frame control);
// after some steps a compaction checkpoint follows (plan, moves, results, new tile flags).
// An item waits only for items listed before it (the previous phase of its own tile, or its
// code's checkpoint), so whichever workgroups run, every wait ends: the items waited for were
// taken earlier by running workgroups, and those wait only for items earlier still.
// FeCl's per-frame schedule (src/detail/BpDecoder/BpDecoderImpl.cpp:49-61) is the same as in
// the per-step launches: the per-tile counters replace kernel boundaries.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ldpc_kernels.h"

namespace fecl {
namespace persist {

enum : int32_t { K_CN = 0, K_VN = 1, K_SYN = 2, K_PLAN = 3, K_MOVE = 4, K_RES = 5, K_RETILE = 6, K_TAIL = 7 };

constexpr int kMaxCodes = 16;
// A compaction checkpoint follows steps 2, 5, 8, ... while at least kMinSteps steps remain.
constexpr int kCpEvery = 3;
constexpr int kMinSteps = 4;
// Work per item: check rows and variables per wavefront, syndrome rows, moved edges and
// variables, result variables.
constexpr int kRowsPerWave = 64;
constexpr int kVarsPerWave = 64;
constexpr int kCnRows = kWavesPerBlock * kRowsPerWave;   // 256 rows per check item
constexpr int kVnVars = kWavesPerBlock * kVarsPerWave;   // 256 variables per variable item
constexpr int kSynRows = 4 * kBlock;                     // 1024 rows per syndrome item
constexpr int kMoveSpan = 256;                           // edges (or variables) per move item
constexpr int kResVars = 8 * kBlock;                     // 2048 variables per results item

// Items of one code in one phase block of a step: tiles * chunks.
struct Block {
  int32_t kind, code, count, chunks;
};

struct Queue {
  int32_t ncodes;
  int32_t imax;                        // the largest iteration count among the codes
  int32_t nblk[2];                     // blocks of a plain step and of a checkpoint step
  Block blk[2][7 * kMaxCodes];
  int64_t per_step[2];                 // items of a plain and of a checkpoint step
  int64_t total;                       // items of all steps
  int32_t tail_off[kMaxCodes + 1];     // final-results items of each code (second queue)
};

__host__ __device__ inline bool is_checkpoint(const Queue& q, int s) {
  return (s + 1) % kCpEvery == 0 && s + kMinSteps < q.imax;
}
// Checkpoint number of checkpoint step s (0 for the first).
__host__ __device__ inline int cp_index(int s) { return (s + 1) / kCpEvery - 1; }
__host__ __device__ inline int64_t step_base(const Queue& q, int s) {
  // checkpoint steps before s: s' = kCpEvery * j - 1 for j >= 1 with s' < s (j <= s / kCpEvery)
  // and s' + kMinSteps < imax (j <= (imax - kMinSteps) / kCpEvery)
  const int64_t lim = q.imax >= kMinSteps ? (q.imax - kMinSteps) / kCpEvery : 0;
  int64_t ncp = s / kCpEvery;
  if (ncp > lim) ncp = lim;
  return (int64_t)s * q.per_step[0] + ncp * (q.per_step[1] - q.per_step[0]);
}

struct Item {
  int32_t kind, code, tile, chunk, step;
};

// Item i (0 <= i < q.total) of the step queue.
__host__ __device__ inline Item decode_item(const Queue& q, int64_t i) {
  int lo = 0, hi = q.imax;   // the step s with step_base(s) <= i < step_base(s + 1)
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (step_base(q, mid) <= i) lo = mid;
    else hi = mid;
  }
  Item it{-1, 0, 0, 0, lo};
  int64_t r = i - step_base(q, lo);
  const int ty = is_checkpoint(q, lo) ? 1 : 0;
  for (int b = 0; b < q.nblk[ty]; ++b) {
    const Block& bl = q.blk[ty][b];
    if (r < bl.count) {
      it.kind = bl.kind;
      it.code = bl.code;
      it.tile = (int32_t)(r / bl.chunks);
      it.chunk = (int32_t)(r % bl.chunks);
      return it;
    }
    r -= bl.count;
  }
  return it;
}

// Per-code chunk counts and sizes the queue is built from.
struct CodeShape {
  int32_t tiles;                       // tiles of the first region (and of each area)
  int32_t iters;                       // the code's iteration count
  int32_t cn_chunks, vn_chunks, syn_chunks, mv_chunks, res_chunks;
};

inline void chunk_counts(int64_t M, int64_t N, int64_t K, int64_t E, CodeShape* cs) {
  cs->cn_chunks = (int32_t)((M + kCnRows - 1) / kCnRows);
  cs->vn_chunks = (int32_t)((N + kVnVars - 1) / kVnVars);
  cs->syn_chunks = (int32_t)((M + kSynRows - 1) / kSynRows);
  // message edges, channel values, then one item of per-slot state
  cs->mv_chunks = (int32_t)((E + kMoveSpan - 1) / kMoveSpan + (N + kMoveSpan - 1) / kMoveSpan + 1);
  cs->res_chunks = (int32_t)((K + kResVars - 1) / kResVars);
}

inline void build_queue(const CodeShape* cs, int n, Queue* q) {
  *q = Queue{};
  q->ncodes = n;
  for (int c = 0; c < n; ++c) q->imax = cs[c].iters > q->imax ? cs[c].iters : q->imax;
  for (int ty = 0; ty < 2; ++ty) {
    int nb = 0;
    int64_t sum = 0;
    auto add = [&](int32_t kind, int c, int32_t count, int32_t chunks) {
      q->blk[ty][nb++] = Block{kind, c, count, chunks};
      sum += count;
    };
    for (int c = 0; c < n; ++c) add(K_CN, c, cs[c].tiles * cs[c].cn_chunks, cs[c].cn_chunks);
    for (int c = 0; c < n; ++c) add(K_VN, c, cs[c].tiles * cs[c].vn_chunks, cs[c].vn_chunks);
    for (int c = 0; c < n; ++c) add(K_SYN, c, cs[c].tiles * cs[c].syn_chunks, cs[c].syn_chunks);
    if (ty == 1) {
      for (int c = 0; c < n; ++c) add(K_PLAN, c, 1, 1);
      for (int c = 0; c < n; ++c) add(K_MOVE, c, cs[c].tiles * cs[c].mv_chunks, cs[c].mv_chunks);
      for (int c = 0; c < n; ++c) add(K_RES, c, cs[c].tiles * cs[c].res_chunks, cs[c].res_chunks);
      for (int c = 0; c < n; ++c) add(K_RETILE, c, cs[c].tiles, 1);
    }
    q->nblk[ty] = nb;
    q->per_step[ty] = sum;
  }
  q->total = step_base(*q, q->imax);
  q->tail_off[0] = 0;
  for (int c = 0; c < n; ++c) q->tail_off[c + 1] = q->tail_off[c] + cs[c].tiles * cs[c].res_chunks;
}

// Counters of one code (device memory, reset before every launch).  Per tile: check,
// variable and syndrome items finished (over all steps), steps whose control finished, and
// whether the tile of the current region holds no live frame.
struct Counters {
  int32_t ctrl_all;     // tile controls finished, over all tiles and steps
  int32_t plan_done;    // checkpoint plans finished
  int32_t res_done, mv_done;   // checkpoint results and move items finished
  int32_t retile_pre;   // checkpoint tile-flag items finished
  int32_t retile_done;  // checkpoints completed (the area flip included)
  int32_t go, new_tiles, live;   // the last plan's decision
  int32_t cur_area, cur_tiles;   // the current region
  int32_t live_tiles;   // tiles of the current region that still hold a live frame
  int32_t done;         // no live frame left in the code
  int32_t pad[3];
};
enum : int32_t { TC_CN = 0, TC_VN = 1, TC_SYN = 2, TC_CTRL = 3, TC_RETIRED = 4, TC_COUNT = 5 };

struct Global {
  int32_t queue, tail, codes_left, all_done, error, pad[11];
};

}  // namespace persist
}  // namespace fecl
