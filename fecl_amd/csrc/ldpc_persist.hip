// Persistent multi-code min-sum decode: one kernel launch walks the work queue of
// ldpc_persist.h for the batches of several codes, with the same row, variable, syndrome,
// control and compaction arithmetic as the per-step launches (ldpc_device.h), so its results
// are bit-identical to theirs.  See ldpc_persist.h for the queue and why every wait ends.
//
// Synchronisation: a workgroup that finishes an item publishes it with a device-scope fence
// and an atomic increment of a counter; a workgroup whose item depends on it polls that
// counter (one lane, with s_sleep between polls), then fences before reading.  Every poll loop
// also ends when the code has no live frame left, when another workgroup reported an error,
// or after kSpinCap polls (a watchdog: the launch then ends and reports the error instead of
// hanging).
//
// Compaction (as the per-step launches do it, decided on the device): after a checkpoint step
// a plan counts the live frames of the code's current region and, when they fit in three
// quarters of its tiles or fewer and the tile-steps saved outweigh moving them, moves them in
// slot order into the first tiles of the code's other area.  The region left behind hands
// its finished frames' results over first.  Areas alternate, each as large as the first
// region, so every frame moves at most once per checkpoint.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "ldpc_device.h"
#include "ldpc_engine.h"
#include "ldpc_persist.h"

namespace fecl {
namespace persist {

// One code of a persistent batch (device memory, read-only during a launch).
struct PCode {
  DevCode c;
  DevOpts o;
  DevState area[2];          // both areas, tiles = T (a region uses the first cur_tiles)
  uint64_t* out_hard[2];     // decode hard outputs of each area
  int32_t* map[2];           // slot -> frame of each area's region (-1: none)
  int32_t* src_of;           // plan: destination slot -> source slot
  uint8_t* msg_bits;         // [frames][K] results
  int32_t* iterations;       // [frames] or null
  int32_t frames, T, iters, max_dc, max_dv, compaction;
  CodeShape shape;
  Counters* k;
  int32_t* tc;               // [TC_COUNT][T] per-tile counters
};

constexpr uint32_t kSpinCap = 1u << 22;

__device__ static inline int32_t ld(const int32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ static inline void st(int32_t* p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ static inline int32_t add(int32_t* p, int32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Thread 0: waits until *p >= target.  Returns false when the code is done (*esc set), when an
// error was reported, or when the watchdog fires (which reports one).
__device__ static bool wait_ge(const int32_t* p, int32_t target, const int32_t* esc, Global* g) {
  for (uint32_t n = 0;; ++n) {
    if (ld(p) >= target) {
      __threadfence();
      return true;
    }
    if (esc && ld(esc)) return false;
    if ((n & 63) == 63 && ld(&g->error)) return false;
    if (n >= kSpinCap) {
      st(&g->error, 1);
      return false;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// Every thread: publish this workgroup's writes, then thread 0 adds 1 to *p; returns the old
// value to thread 0 (0 elsewhere).
__device__ static inline int32_t publish(int32_t* p) {
  __syncthreads();
  int32_t old = 0;
  if (threadIdx.x == 0) {
    __threadfence();
    old = add(p, 1);
  }
  return old;
}

__device__ static inline int degree_class(int d) { return d <= 4 ? 4 : d <= 8 ? 8 : d <= 16 ? 16 : d <= 32 ? 32 : 0; }

// The phase bodies below are compiled as separate functions (noinline), each for one degree
// bound, so the persistent kernel needs the registers of its heaviest phase rather than of all
// of them inlined together.  They take the code's descriptor by address and read what they
// need from it.

// The region of area a with n tiles.
__device__ static inline DevState region(const PCode* P, int a, int n) {
  DevState s = P->area[a];
  s.tiles = n;
  return s;
}

// Check rows [r0, r0 + kCnRows) of tile t: wavefront w takes kRowsPerWave consecutive rows.
template <class F, int D>
__device__ static inline void cn_rows(const PCode* P, int a, int n, int t, int r0) {
  const DevState s = region(P, a, n);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rb = r0 + wave * kRowsPerWave;
#pragma nounroll
  for (int j = 0; j < kRowsPerWave; ++j) {
    const int r = __builtin_amdgcn_readfirstlane(rb + j);
    if (r >= P->c.M) break;
    cn_minsum_raw_row<F, D, false>(P->c, s, P->o, t, r, lane);
  }
}

template <class F>
__device__ static void run_cn(const PCode* P, int a, int n, int t, int r0) {
  const int dc = degree_class(P->max_dc);
  if (dc <= 8) {
    cn_rows<F, 8>(P, a, n, t, r0);
  } else if (dc == 16) {
    if constexpr (16 * F::W <= 64) cn_rows<F, 16>(P, a, n, t, r0);
  } else {
    if constexpr (32 * F::W <= 64) cn_rows<F, 32>(P, a, n, t, r0);
  }
}

// Variables [v0, v0 + kVnVars) of tile t.
template <class F, int VMAX>
__device__ static inline void vn_rows(const PCode* P, int a, int n, int t, int v0) {
  const DevState s = region(P, a, n);
  DevOut out{};
  out.hard = P->out_hard[a];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int vb = v0 + wave * kVarsPerWave;
#pragma nounroll
  for (int j = 0; j < kVarsPerWave; ++j) {
    const int v = __builtin_amdgcn_readfirstlane(vb + j);
    if (v >= P->c.N) break;
    vn_vars<F, VMAX, 1, false>(P->c, s, out, 0, 0, t, v, lane);
  }
}

template <class F>
__device__ static void run_vn(const PCode* P, int a, int n, int t, int v0) {
  constexpr int FPL = F::FPL;
  const int vs = P->c.vstride;
  if (vs == 4) {
    vn_rows<F, 4>(P, a, n, t, v0);
  } else if (vs == 8) {
    vn_rows<F, 8>(P, a, n, t, v0);
  } else if (vs == 16) {
    if constexpr (16 * FPL <= 64) vn_rows<F, 16>(P, a, n, t, v0);
  } else {
    if constexpr (32 * FPL <= 64) vn_rows<F, 32>(P, a, n, t, v0);
  }
}

template <int FPL, int D>
__device__ static inline void syn_rows(const PCode* P, int a, int n, int t, int r0,
                                                   uint64_t (*part)[FPL]) {
  const DevState s = region(P, a, n);
#pragma nounroll
  for (int b = 0; b < kSynRows / kBlock; ++b) {
    const int rb = r0 + b * kBlock;
    if (rb >= P->c.M) break;
    syndrome_rows<FPL, D>(P->c, s, t, rb, part);
  }
}

template <int FPL>
__device__ static void run_syn(const PCode* P, int a, int n, int t, int r0, uint64_t (*part)[FPL]) {
  const int dc = degree_class(P->max_dc);
  if (dc <= 8) syn_rows<FPL, 8>(P, a, n, t, r0, part);
  else if (dc == 16) syn_rows<FPL, 16>(P, a, n, t, r0, part);
  else if (dc == 32) syn_rows<FPL, 32>(P, a, n, t, r0, part);
  else syn_rows<FPL, 0>(P, a, n, t, r0, part);
}

// Control of tile t (wavefront 0); returns whether a frame is live.
template <int FPL>
__device__ static inline bool ctrl_tile(const PCode* P, int a, int n, int t, int lane) {
  const DevState s = region(P, a, n);
  return control_tile<FPL>(s, P->iters, 0, -1, t, lane);
}

template <int FPL>
__device__ static inline void retile_one(const PCode* P, int area, int t, int lane) {
  (void)retile_tile<FPL>(P->area[area], t, lane);
}

// Plan of a checkpoint (whole workgroup): live slots of the region (area a, n tiles) in slot
// order become destination slots 0 .. live-1 of the other area.  Decides as the per-step
// launches do: go when the live frames fit in three quarters of the tiles or fewer and twice
// the tile-steps freed over the remaining steps exceed three times the live frames (a move
// costs about 1.5 frame-iterations).
template <int TW>
__device__ static inline void run_plan(const PCode* P, int a, int n, int remaining, int32_t* sums) {
  const uint8_t* phase = P->area[a].phase;
  const int slots = n * TW;
  const int tid = threadIdx.x;
  const int per = (slots + kBlock - 1) / kBlock;
  const int i0 = min(slots, tid * per), i1 = min(slots, i0 + per);
  int cnt = 0;
  for (int i = i0; i < i1; ++i) {
    const int b = phase[i] & PH_MASK;
    cnt += (b == PH_RUN || b == PH_FINAL) ? 1 : 0;
  }
  sums[tid] = cnt;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {   // inclusive scan
    const int v = tid >= off ? sums[tid - off] : 0;
    __syncthreads();
    sums[tid] += v;
    __syncthreads();
  }
  const int live = sums[kBlock - 1];
  const int need = (live + TW - 1) / TW;
  const bool go = P->compaction && live > 0 && 4 * need <= 3 * n &&
                  2 * (int64_t)(n - need) * remaining * TW > 3 * (int64_t)live;
  if (go) {
    const int32_t* map_in = P->map[a];
    int32_t* map_out = P->map[1 - a];
    int j = sums[tid] - cnt;
    for (int i = i0; i < i1; ++i) {
      const int b = phase[i] & PH_MASK;
      if (b == PH_RUN || b == PH_FINAL) {
        P->src_of[j] = i;
        map_out[j] = map_in[i];
        ++j;
      }
    }
    for (int d = live + tid; d < need * TW; d += kBlock) {
      P->src_of[d] = -1;
      map_out[d] = -1;
    }
  }
  if (tid == 0) {
    st(&P->k->go, go ? 1 : 0);
    st(&P->k->new_tiles, go ? need : n);
    st(&P->k->live, live);
  }
}

// Results of the frames that finished in tile t of the region in area a: hard bits of
// variables [v0, v0 + kResVars) (and, for the first chunk, iteration counts) of every DONE slot.
template <int FPL>
__device__ static inline void run_results(const PCode* P, int a, int t, int chunk) {
  constexpr int TW = kWave * FPL;
  const uint8_t* phase = P->area[a].phase;
  const int32_t* iter = P->area[a].iter;
  const int32_t* map = P->map[a];
  const int K = P->c.K;
  const int v0 = chunk * kResVars;
  const uint64_t* hard = P->out_hard[a];
  for (int p = 0; p < TW; ++p) {
    const int slot = t * TW + p;
    const int f = map[slot];
    if (f < 0 || f >= P->frames || (phase[slot] & PH_MASK) != PH_DONE) continue;
    for (int v = v0 + threadIdx.x; v < K && v < v0 + kResVars; v += kBlock) {
      const uint64_t w = hard[((int64_t)t * K + v) * FPL + (p % FPL)];
      P->msg_bits[(int64_t)f * K + v] = (uint8_t)((w >> (p / FPL)) & 1ull);
    }
  }
  if (chunk == 0 && P->iterations) {
    for (int p = threadIdx.x; p < TW; p += kBlock) {
      const int slot = t * TW + p;
      const int f = map[slot];
      if (f >= 0 && f < P->frames && (phase[slot] & PH_MASK) == PH_DONE) P->iterations[f] = iter[slot];
    }
  }
}

// Move item `chunk` of destination tile t from area a (the old region) into area 1 - a.
template <class F>
__device__ static inline void run_move(const PCode* P, int a, int t, int chunk) {
  constexpr int TW = F::TW;
  const DevState& from = P->area[a];
  const DevState& to = P->area[1 - a];
  const int msg_chunks = (P->c.E + kMoveSpan - 1) / kMoveSpan;
  const int l_chunks = (P->c.N + kMoveSpan - 1) / kMoveSpan;
  if (chunk < msg_chunks) {
    for (int e = chunk * kMoveSpan; e < (chunk + 1) * kMoveSpan && e < P->c.E; e += kCompactEdges)
      compact_msg_block<F>(P->c, from.msg, to.msg, P->src_of, t, e);
  } else if (chunk < msg_chunks + l_chunks) {
    const int vb = (chunk - msg_chunks) * kMoveSpan;
    for (int v = vb; v < vb + kMoveSpan && v < P->c.N; v += kCompactEdges)
      compact_L_block<TW>(P->c, from.L, to.L, P->src_of, t, v);
  } else {
    for (int j = t * TW + threadIdx.x; j < (t + 1) * TW; j += kBlock)
      compact_slot(from.phase, from.iter, to.phase, to.iter, P->src_of, j);
  }
}

template <class F>
__global__ __launch_bounds__(kBlock) void persist_kernel(const PCode* __restrict__ codes, const Queue* __restrict__ qp,
                                                         Global* g) {
  constexpr int FPL = F::FPL, TW = F::TW;
  __shared__ int32_t info[12];
  __shared__ uint64_t part[kWavesPerBlock][FPL];
  __shared__ int32_t sums[kBlock];
  const Queue& q = *qp;
  const int tid = threadIdx.x;
  // ---- the steps
  for (;;) {
    if (tid == 0) {
      info[0] = -1;
      if (!ld(&g->all_done) && !ld(&g->error)) {
        const int64_t i = add(&g->queue, 1);
        if (i < q.total) {
          const Item it = decode_item(q, i);
          info[0] = it.kind;
          info[1] = it.code;
          info[2] = it.tile;
          info[3] = it.chunk;
          info[4] = it.step;
        }
      }
    }
    __syncthreads();
    // item fields from LDS, made wave-uniform so that descriptors and addresses stay scalar
    const int kind = __builtin_amdgcn_readfirstlane(info[0]);
    if (kind < 0) break;
    const int c = __builtin_amdgcn_readfirstlane(info[1]), t = __builtin_amdgcn_readfirstlane(info[2]);
    const int chunk = __builtin_amdgcn_readfirstlane(info[3]), s = __builtin_amdgcn_readfirstlane(info[4]);
    const PCode* Pp = codes + c;
    const PCode& P = *Pp;
    Counters* k = P.k;
    int32_t* tc_cn = P.tc + TC_CN * P.T;
    int32_t* tc_vn = P.tc + TC_VN * P.T;
    int32_t* tc_syn = P.tc + TC_SYN * P.T;
    int32_t* tc_ctrl = P.tc + TC_CTRL * P.T;
    int32_t* tc_ret = P.tc + TC_RETIRED * P.T;
    // thread 0: wait for the item's dependency; info[5] = run (1) / skip but count (0) /
    // skip without counting (-1: the code is done or an error was reported)
    if (tid == 0) {
      int run = -1;
      bool ok = false;
      if (!ld(&k->done)) {
        switch (kind) {
          case K_CN:
            if (s == 0) ok = true;
            else if (is_checkpoint(q, s - 1)) ok = wait_ge(&k->retile_done, cp_index(s - 1) + 1, &k->done, g);
            else ok = wait_ge(&tc_ctrl[t], s, &k->done, g);
            break;
          case K_VN: ok = wait_ge(&tc_cn[t], (s + 1) * P.shape.cn_chunks, &k->done, g); break;
          case K_SYN: ok = wait_ge(&tc_vn[t], (s + 1) * P.shape.vn_chunks, &k->done, g); break;
          case K_PLAN: ok = wait_ge(&k->ctrl_all, (s + 1) * P.T, &k->done, g); break;
          case K_MOVE:
          case K_RES: ok = wait_ge(&k->plan_done, cp_index(s) + 1, &k->done, g); break;
          case K_RETILE:
            ok = wait_ge(&k->mv_done, (cp_index(s) + 1) * P.T * P.shape.mv_chunks, &k->done, g) &&
                 wait_ge(&k->res_done, (cp_index(s) + 1) * P.T * P.shape.res_chunks, &k->done, g);
            break;
        }
      }
      if (ok) {
        const int a = ld(&k->cur_area), n = ld(&k->cur_tiles);
        info[6] = a;
        info[7] = n;
        switch (kind) {
          case K_CN:
          case K_VN:
          case K_SYN: run = (t < n && !ld(&tc_ret[t]) && s < P.iters) ? 1 : 0; break;
          case K_PLAN: run = (P.compaction && s + kMinSteps < P.iters) ? 1 : 0; break;
          case K_MOVE: run = (ld(&k->go) && t < ld(&k->new_tiles)) ? 1 : 0; break;
          case K_RES: run = (ld(&k->go) && t < n) ? 1 : 0; break;
          case K_RETILE: run = 1; break;
        }
      }
      info[5] = run;
    }
    __syncthreads();
    const int run = __builtin_amdgcn_readfirstlane(info[5]);
    if (run < 0) {
      __syncthreads();
      continue;
    }
    const int a = __builtin_amdgcn_readfirstlane(info[6]), n = __builtin_amdgcn_readfirstlane(info[7]);
    switch (kind) {
      case K_CN:
        if (run) run_cn<F>(Pp, a, n, t, chunk * kCnRows);
        publish(&tc_cn[t]);
        break;
      case K_VN:
        if (run) run_vn<F>(Pp, a, n, t, chunk * kVnVars);
        publish(&tc_vn[t]);
        break;
      case K_SYN: {
        if (run) run_syn<FPL>(Pp, a, n, t, chunk * kSynRows, part);
        const int32_t old = publish(&tc_syn[t]);
        // the workgroup finishing the tile's last syndrome item of step s runs its control
        if (tid == 0) info[8] = (old + 1 == (s + 1) * P.shape.syn_chunks) ? 1 : 0;
        __syncthreads();
        if (__builtin_amdgcn_readfirstlane(info[8])) {
          if (tid == 0) __threadfence();
          __syncthreads();
          if (run && tid < 64) {
            const bool live = ctrl_tile<FPL>(Pp, a, n, t, tid);
            if (tid == 0 && !live) {
              st(&tc_ret[t], 1);
              if (add(&k->live_tiles, -1) == 1) {
                st(&k->done, 1);
                if (add(&g->codes_left, -1) == 1) st(&g->all_done, 1);
              }
            }
          }
          __syncthreads();
          if (tid == 0) {
            __threadfence();
            add(&tc_ctrl[t], 1);
            add(&k->ctrl_all, 1);
          }
        }
        break;
      }
      case K_PLAN:
        if (run) {
          run_plan<TW>(Pp, a, n, P.iters - 1 - s, sums);
        } else if (tid == 0) {
          st(&k->go, 0);
          st(&k->new_tiles, n);
        }
        publish(&k->plan_done);
        break;
      case K_MOVE:
        if (run) run_move<F>(Pp, a, t, chunk);
        publish(&k->mv_done);
        break;
      case K_RES:
        if (run) run_results<FPL>(Pp, a, t, chunk);
        publish(&k->res_done);
        break;
      case K_RETILE: {
        if (tid == 0) {
          info[9] = ld(&k->go);
          info[10] = ld(&k->new_tiles);
        }
        __syncthreads();
        const int go = __builtin_amdgcn_readfirstlane(info[9]), nt = __builtin_amdgcn_readfirstlane(info[10]);
        if (go) {
          if (t < nt) {
            if (tid < 64) retile_one<FPL>(Pp, 1 - a, t, tid);
            if (tid == 0) st(&tc_ret[t], 0);
          } else if (tid == 0) {
            st(&tc_ret[t], 1);
          }
        }
        const int32_t old = publish(&k->retile_pre);
        if (tid == 0 && old + 1 == (cp_index(s) + 1) * P.T) {
          // the checkpoint's last item: switch to the new region, then release it
          if (go) {
            st(&k->cur_area, 1 - a);
            st(&k->cur_tiles, nt);
            st(&k->live_tiles, nt);
          }
          __threadfence();
          st(&k->retile_done, cp_index(s) + 1);
        }
        break;
      }
    }
    __syncthreads();
  }
  // ---- final results of every code's last region
  const int ncodes = q.ncodes;
  for (;;) {
    if (tid == 0) {
      info[0] = -1;
      if (!ld(&g->error)) {
        const int i = add(&g->tail, 1);
        if (i < q.tail_off[ncodes]) {
          int c = 0;
          while (i >= q.tail_off[c + 1]) ++c;
          const int r = i - q.tail_off[c];
          info[0] = K_TAIL;
          info[1] = c;
          info[2] = r / codes[c].shape.res_chunks;
          info[3] = r % codes[c].shape.res_chunks;
          const Counters* k = codes[c].k;
          info[5] = wait_ge(&k->done, 1, nullptr, g) ? 1 : 0;
          info[6] = ld(&k->cur_area);
          info[7] = ld(&k->cur_tiles);
        }
      }
    }
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(info[0]) < 0) break;
    const int c = __builtin_amdgcn_readfirstlane(info[1]), t = __builtin_amdgcn_readfirstlane(info[2]);
    const int chunk = __builtin_amdgcn_readfirstlane(info[3]);
    if (__builtin_amdgcn_readfirstlane(info[5]) && t < __builtin_amdgcn_readfirstlane(info[7]))
      run_results<FPL>(codes + c, __builtin_amdgcn_readfirstlane(info[6]), t, chunk);
    __syncthreads();
  }
}

// Per-launch set-up after the counters were zeroed: the first area's slot map (slot j holds
// frame j) and each code's current region.
__global__ void persist_reset_kernel(const PCode* codes, int ncodes, Global* g, int32_t tw) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int c = 0; c < ncodes; ++c) {
    const PCode& P = codes[c];
    const int64_t slots = (int64_t)P.T * tw;
    if (i < slots) P.map[0][i] = i < P.frames ? (int32_t)i : -1;
    if (i == 0) {
      P.k->cur_area = 0;
      P.k->cur_tiles = P.T;
      P.k->live_tiles = P.T;
    }
  }
  if (i == 0) g->codes_left = ncodes;
}

}  // namespace persist

// ------------------------------------------------------------------------------------------
// Host side.

using namespace persist;

namespace engine {

struct PersistBatch {
  Layout lay;
  int ncodes = 0;
  std::vector<std::unique_ptr<Workspace>> ws;
  DevBuf ctl;             // Global | Queue | PCode[n] | Counters[n] | per-tile counters
  size_t zero_off = 0, zero_words = 0;
  Global* g = nullptr;
  Queue* q = nullptr;
  PCode* codes = nullptr;
  int grid = 0;
  std::vector<int64_t> frames;
  std::vector<const float*> llr;
  std::vector<DevState> area0;   // each code's first region, for its ingest
  ~PersistBatch() { ctl.release(); }
};

void persist_destroy(PersistBatch* b) { delete b; }

// Whether the handles can share one persistent launch: min-sum without scaling factors (raw
// encodings), one message format and lane layout whose check rows are register-staged, no
// fused syndrome, one device, at most kMaxCodes codes.
bool persist_eligible(fecl_ldpc* const* handles, int32_t count, Layout* common) {
  if (count < 1 || count > kMaxCodes) return false;
  Layout lay = layout_of(handles[0]);
  lay.vpw = 1;
  for (int32_t i = 0; i < count; ++i) {
    const fecl_ldpc* h = handles[i];
    const Layout li = layout_of(h);
    if (h->algorithm != FECL_APPROXIMATE || h->sf_mode != 0 || h->fused_syndrome || !h->persistent) return false;
    if (li.fmt != lay.fmt || li.words != lay.words || h->device != handles[0]->device) return false;
    if (!cn_register_staged(lay, h->algorithm, h->sf_mode, h->code.max_row_degree)) return false;
    if (!layout_fits(lay, h->algorithm, h->code.max_row_degree, h->code.max_var_degree)) return false;
    if (h->code.max_var_degree > 32) return false;
  }
  if (lay.words != 1 && lay.words != 2 && lay.words != 4) return false;
  *common = lay;
  return true;
}

template <class F>
static hipError_t launch_persist(PersistBatch* b, hipStream_t st) {
  hipLaunchKernelGGL((persist_kernel<F>), dim3(b->grid), dim3(kBlock), 0, st, b->codes, b->q, b->g);
  return hipGetLastError();
}

template <class F>
static int grid_for() {
  int dev = 0, blocks = 0;
  hipDeviceProp_t prop{};
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, persist_kernel<F>, kBlock, 0) != hipSuccess) return 0;
  return std::max(1, blocks) * prop.multiProcessorCount;
}

#define FECL_PERSIST_DISPATCH(lay, CALL)                                                        \
  switch (lay.fmt * 8 + lay.words) {                                                           \
    case 0 * 8 + 1: { using F = Fmt<ElemF32, 1>; CALL; break; }                                 \
    case 0 * 8 + 2: { using F = Fmt<ElemF32, 2>; CALL; break; }                                 \
    case 0 * 8 + 4: { using F = Fmt<ElemF32, 4>; CALL; break; }                                 \
    case 1 * 8 + 1: { using F = Fmt<ElemF16, 1>; CALL; break; }                                 \
    case 1 * 8 + 2: { using F = Fmt<ElemF16, 2>; CALL; break; }                                 \
    case 1 * 8 + 4: { using F = Fmt<ElemF16, 4>; CALL; break; }                                 \
    case 2 * 8 + 1: { using F = Fmt<ElemBF16, 1>; CALL; break; }                                \
    case 2 * 8 + 2: { using F = Fmt<ElemBF16, 2>; CALL; break; }                                \
    case 2 * 8 + 4: { using F = Fmt<ElemBF16, 4>; CALL; break; }                                \
    default: break;                                                                             \
  }

int persist_create(fecl_ldpc* const* handles, const int64_t* frames, int32_t count, const float* const* llr,
                   uint8_t* const* msg, int32_t* const* iterations, const Layout& lay, const DevOpts* opts,
                   PersistBatch** out) {
  std::unique_ptr<PersistBatch> b(new PersistBatch());
  b->lay = lay;
  b->ncodes = count;
  const int TW = lay.tw();
  std::vector<CodeShape> shapes(count);
  std::vector<PCode> host(count);
  for (int32_t i = 0; i < count; ++i) {
    fecl_ldpc* h = handles[i];
    const Code& c = h->code;
    const int64_t T = (frames[i] + TW - 1) / TW;
    if (T * TW > INT32_MAX / 2) return fail(FECL_INVALID_ARGUMENT, "batch too large for one persistent launch");
    std::unique_ptr<Workspace> ws(new Workspace());
    DevState s0;
    DevOut o0;
    // area 0: the first region; area 1 (ws->spare): as many tiles; slot maps and plan scratch
    if (int rc = prepare(h, ws.get(), lay, frames[i], false, &s0, &o0)) return rc;
    {
      Area& x = ws->spare;
      HIP_TRY(x.msg.ensure((size_t)T * c.E * lay.rw() * 4));
      HIP_TRY(x.L.ensure((size_t)T * c.N * TW * 4));
      HIP_TRY(x.hard.ensure((size_t)T * c.N * lay.fpl() * 8));
      HIP_TRY(x.phase.ensure((size_t)T * TW));
      HIP_TRY(x.iter.ensure((size_t)T * TW * 4));
      HIP_TRY(x.fail.ensure((size_t)T * lay.fpl() * 8));
      HIP_TRY(x.flags.ensure((size_t)T * 8));
      HIP_TRY(x.out_hard.ensure((size_t)T * c.K * lay.fpl() * 8));
      for (DevBuf* m : {&ws->map[0], &ws->map[1], &ws->src_of}) HIP_TRY(m->ensure((size_t)T * TW * 4));
    }
    CodeShape& cs = shapes[i];
    cs.tiles = (int32_t)T;
    cs.iters = h->iterations;
    chunk_counts(c.M, c.N, c.K, c.E, &cs);
    PCode& P = host[i];
    P = PCode{};
    P.c = h->dc;
    P.o = opts[i];
    P.o.nt_rows = 0;
    P.o.fused_syn = 0;
    P.area[0] = s0;
    P.area[0].active = nullptr;
    DevState s1 = s0;
    s1.msg = ws->spare.msg.as<uint32_t>();
    s1.L = ws->spare.L.as<float>();
    s1.hard = ws->spare.hard.as<uint64_t>();
    s1.phase = ws->spare.phase.as<uint8_t>();
    s1.iter = ws->spare.iter.as<int32_t>();
    s1.fail = ws->spare.fail.as<uint64_t>();
    s1.tile_flags = ws->spare.flags.as<uint32_t>();
    s1.tile_iter = ws->spare.flags.as<int32_t>() + T;
    s1.active = nullptr;
    P.area[1] = s1;
    P.out_hard[0] = o0.hard;
    P.out_hard[1] = ws->spare.out_hard.as<uint64_t>();
    P.map[0] = ws->map[0].as<int32_t>();
    P.map[1] = ws->map[1].as<int32_t>();
    P.src_of = ws->src_of.as<int32_t>();
    P.msg_bits = msg[i];
    P.iterations = iterations ? iterations[i] : nullptr;
    P.frames = (int32_t)frames[i];
    P.T = (int32_t)T;
    P.iters = h->iterations;
    P.max_dc = c.max_row_degree;
    P.max_dv = c.max_var_degree;
    P.compaction = h->compaction ? 1 : 0;
    P.shape = cs;
    b->ws.push_back(std::move(ws));
    b->frames.push_back(frames[i]);
    b->llr.push_back(llr[i]);
    b->area0.push_back(s0);
  }
  // control block: Global, Queue, PCode[n], then the zeroed counters
  auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t off_q = align(sizeof(Global));
  const size_t off_codes = off_q + align(sizeof(Queue));
  const size_t off_ctr = off_codes + align(sizeof(PCode) * count);
  size_t tiles_total = 0;
  for (auto& cs : shapes) tiles_total += (size_t)cs.tiles;
  const size_t ctr_bytes = sizeof(Counters) * count + sizeof(int32_t) * TC_COUNT * tiles_total;
  const size_t bytes = off_ctr + align(ctr_bytes);
  HIP_TRY(b->ctl.ensure(bytes));
  char* base = b->ctl.as<char>();
  b->g = reinterpret_cast<Global*>(base);
  b->q = reinterpret_cast<Queue*>(base + off_q);
  b->codes = reinterpret_cast<PCode*>(base + off_codes);
  Counters* ctrs = reinterpret_cast<Counters*>(base + off_ctr);
  int32_t* tcs = reinterpret_cast<int32_t*>(base + off_ctr + sizeof(Counters) * count);
  size_t tpos = 0;
  for (int32_t i = 0; i < count; ++i) {
    host[i].k = ctrs + i;
    host[i].tc = tcs + tpos;
    tpos += (size_t)TC_COUNT * shapes[i].tiles;
  }
  Queue qh;
  build_queue(shapes.data(), count, &qh);
  std::vector<char> blob(off_ctr, 0);
  std::memcpy(blob.data() + off_q, &qh, sizeof(Queue));
  std::memcpy(blob.data() + off_codes, host.data(), sizeof(PCode) * count);
  HIP_TRY(hipMemcpy(base, blob.data(), off_ctr, hipMemcpyHostToDevice));
  b->zero_off = off_ctr;
  b->zero_words = ctr_bytes / 4;
  FECL_PERSIST_DISPATCH(lay, b->grid = grid_for<F>());
  if (b->grid <= 0) return fail(FECL_DEVICE_ERROR, "persistent kernel occupancy query failed");
  *out = b.release();
  return FECL_OK;
}

// Enqueues one decode of the batch: counter reset, each code's ingest, the persistent kernel.
// The counters are cleared with a fill kernel (no memset node in a captured graph).
int persist_enqueue(PersistBatch* b, fecl_ldpc* const* handles, hipStream_t st) {
  const Layout& lay = b->lay;
  char* base = b->ctl.as<char>();
  HIP_TRY(kernels_fill_i32(reinterpret_cast<int32_t*>(base), (int64_t)(sizeof(Global) / 4), 0, st));
  HIP_TRY(kernels_fill_i32(reinterpret_cast<int32_t*>(base + b->zero_off), (int64_t)b->zero_words, 0, st));
  int64_t max_slots = 1;
  for (int i = 0; i < b->ncodes; ++i)
    max_slots = std::max<int64_t>(max_slots, (b->frames[i] + lay.tw() - 1) / lay.tw() * lay.tw());
  hipLaunchKernelGGL(persist_reset_kernel, dim3((unsigned)((max_slots + 255) / 256)), dim3(256), 0, st, b->codes,
                     b->ncodes, b->g, (int32_t)lay.tw());
  HIP_TRY(hipGetLastError());
  for (int i = 0; i < b->ncodes; ++i) {
    fecl_ldpc* h = handles[i];
    const uint8_t init = (uint8_t)((h->iterations == 1 ? PH_FINAL : PH_RUN) | PH_FRESH);
    HIP_TRY(kernels_ingest(lay, h->dc, b->area0[i], b->llr[i], nullptr, false, b->frames[i], 0, init, st));
  }
  hipError_t e = hipErrorInvalidValue;
  FECL_PERSIST_DISPATCH(lay, e = launch_persist<F>(b, st));
  HIP_TRY(e);
  return FECL_OK;
}

// Whether the last completed launch ended on its watchdog (synchronous read).
int persist_error(PersistBatch* b, int32_t* error) {
  Global g{};
  HIP_TRY(hipMemcpy(&g, b->g, sizeof(Global), hipMemcpyDeviceToHost));
  *error = g.error;
  return FECL_OK;
}

}  // namespace engine
}  // namespace fecl
