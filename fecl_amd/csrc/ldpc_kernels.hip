// HIP kernels of the flooding BP decoder for gfx950 (wave64).  Compiled with
// -ffp-contract=off: every float operation rounds exactly as FeCl's CPU decoder instantiated
// with a float LLR type does, which is what makes the FP32 min-sum and Linear paths
// bit-identical to it.
//
// Reference arithmetic (FeCl):
//   check update     src/detail/BpDecoder/BpDecoderImpl.cpp:133-159
//   bit update       src/detail/BpDecoder/BpDecoderImpl.cpp:162-179
//   schedule/output  src/detail/BpDecoder/BpDecoderImpl.cpp:38-130
//   MinBoxSum / LinearBoxSum / BoxSum   src/detail/LlrMetrics.h:126-196
//   linear table     src/detail/LinearTable.h:57-81
//   syndrome test    src/detail/Ldpc.cpp:160-172
//
// Message rows.  A (tile, edge) row holds the edge's message for the TW frames of the tile,
// packed PW per 32-bit word (PW = 1 for FP32, 2 for FP16/BF16): RW = TW / PW words.  Each lane
// owns W consecutive words (FPL = W * PW frames), so a wavefront moves 256*W contiguous bytes
// per edge; W is the layout knob the autotuner picks (frames per workgroup = 4 * 64 * FPL).
//
// The row, variable, syndrome, control and compaction bodies live in ldpc_device.h as functions
// of their (tile, row / variable), shared with the persistent kernel (ldpc_persist.hip); the
// kernels here map the launch grid onto them.
#include <algorithm>

#include "ldpc_device.h"
#include "ldpc_kernels.h"
#include "ldpc_launch.h"

namespace fecl {

template <class F, int DMAX>
__global__ __launch_bounds__(kBlock) void cn_minsum_kernel(DevCode c, DevState s, DevOpts o) {
  const int r = uniform(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  if (r >= c.M) return;
  cn_minsum_row<F, DMAX>(c, s, o, blockIdx.y, r, threadIdx.x & 63);
}

template <class F, int DMAX, bool NT>
__global__ __launch_bounds__(kBlock) void cn_minsum_raw_kernel(DevCode c, DevState s, DevOpts o) {
  const int r = uniform(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  if (r >= c.M) return;
  cn_minsum_raw_row<F, DMAX, NT>(c, s, o, blockIdx.y, r, threadIdx.x & 63);
}

// Check-node update for Linear (ALG 1) and Exact (ALG 2 here) in FeCl's forward/backward
// order: prefix P_j = ((x_0 # x_1) # ...) # x_{j-1}; suffix from x_{d-1} downwards;
// out_{d-1} = P_{d-1}, out_j = P_j # S_j, out_0 = S_0, each times the scaling factor.  Linear
// combines LLRs pairwise; Exact combines the complements q of its priors the same way and
// applies the post to each output.
template <class F, int DMAX, int ALG>
__global__ __launch_bounds__(kBlock) void cn_box_kernel(DevCode c, DevState s, DevOpts o) {
  constexpr int FPL = F::FPL, TW = F::TW, RW = F::RW;
  __shared__ float2 tab[8];
  if (ALG == 1) {   // before any wave of the workgroup can leave
    if (threadIdx.x < 8) {
      float y = o.lin_y[0], sl = o.lin_s[0];
#pragma unroll
      for (int q = 1; q < 8; ++q)
        if ((int)threadIdx.x == q) { y = o.lin_y[q]; sl = o.lin_s[q]; }
      tab[threadIdx.x] = make_float2(y, sl);
    }
    __syncthreads();
  }
  const int t = blockIdx.y;
  const int r = uniform(blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6));
  if (r >= c.M) return;
  const int lane = threadIdx.x & 63;
  // the row's extent and the tile's flags are requested together (the exit test reads both,
  // which keeps the compiler from deferring the extent load past it)
  const int e0 = uniform(c.row_ptr[r]);
  const int d = uniform(c.row_ptr[r + 1]) - e0;
  const uint32_t tflags = uniform((int)s.tile_flags[t]);
  if (!(tflags & TF_LIVE) || d < 0) return;
  const int ll = lane_for_loads(tflags, lane);
  uint32_t* mrow = s.msg + ((int64_t)t * c.E + e0) * RW + lane * F::W;   // stores
  const uint32_t* mld = mrow + (ll - lane) * F::W;                        // row loads
  const float* Lt = s.L + (int64_t)t * c.N * TW + ll * FPL;
  uint8_t ph[FPL];
  int it[FPL];
  lane_state<FPL>(s, t, tflags, lane, ph, it);
  float x[DMAX][FPL];
#pragma unroll
  for (int j = 0; j < DMAX; ++j)
    if (j < d) F::load(mld + (int64_t)j * RW, x[j]);
  CnLanes<FPL> ln;
  if (!cn_decide<FPL>(o, ph, it, ln)) return;
  if (o.fused_syn && ln.wave_seen) cn_syndrome<FPL>(c, s, t, e0, d, lane);
  if (ln.wave_fresh) {   // frames whose v2c is still L[col(e)]
#pragma unroll
    for (int j = 0; j < DMAX; ++j) {
      if (j < d) {
        const int v = uniform(c.col[e0 + j]);
        float l[FPL];
        load_floats<FPL>(Lt + (int64_t)v * TW, l);
#pragma unroll
        for (int k = 0; k < FPL; ++k)
          if (ln.fresh[k]) x[j][k] = F::quantize(l[k]);
      }
    }
  }
  if (!ln.lane_any) return;
  auto store = [&](int j, float (&y)[FPL]) {
#pragma unroll
    for (int k = 0; k < FPL; ++k) {
      y[k] = F::quantize(apply_sf(o, ln.sfrow[k], d - 2, y[k]));
      if (!ln.act[k]) y[k] = x[j][k];   // other frames keep their word
    }
    F::store(mrow + (int64_t)j * RW, y);
  };
  float pre[DMAX][FPL];
  float P[FPL];
  if constexpr (ALG == 2) {
    // q of every input, and the sign of the product of all priors (each prior has the sign
    // of -x; the post flips it again): out_j's sign bit is (d & 1) ^ sx ^ sign(x_j)
    float q[DMAX][FPL];
    uint32_t sx[FPL];
#pragma unroll
    for (int k = 0; k < FPL; ++k) sx[k] = (d & 1) ? 0x80000000u : 0u;
#pragma unroll
    for (int j = 0; j < DMAX; ++j) {
      if (j < d) {
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
          q[j][k] = exact_prior(x[j][k]);
          sx[k] ^= __float_as_uint(x[j][k]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < FPL; ++k) P[k] = q[0][k];
#pragma unroll
    for (int j = 1; j < DMAX - 1; ++j) {
      if (j < d - 1) {
#pragma unroll
        for (int k = 0; k < FPL; ++k) { pre[j][k] = P[k]; P[k] = exact_comb(P[k], q[j][k]); }
      }
    }
#pragma unroll
    for (int j = DMAX - 1; j >= 1; --j) {
      if (j == d - 1) {
        float y[FPL];
#pragma unroll
        for (int k = 0; k < FPL; ++k) y[k] = exact_post(P[k], sx[k] ^ __float_as_uint(x[j][k]));
        store(j, y);
      }
    }
    float S[FPL];
#pragma unroll
    for (int j = DMAX - 1; j >= 1; --j) {
      if (j == d - 1) {
#pragma unroll
        for (int k = 0; k < FPL; ++k) S[k] = q[j][k];
      }
    }
#pragma unroll
    for (int j = DMAX - 2; j >= 1; --j) {
      if (j <= d - 2) {
        float y[FPL];
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
          y[k] = exact_post(exact_comb(pre[j][k], S[k]), sx[k] ^ __float_as_uint(x[j][k]));
          S[k] = exact_comb(S[k], q[j][k]);
        }
        store(j, y);
      }
    }
    float y[FPL];
#pragma unroll
    for (int k = 0; k < FPL; ++k) y[k] = exact_post(S[k], sx[k] ^ __float_as_uint(x[0][k]));
    store(0, y);
  } else {
    // forward prefixes: pre[j] for j = 1 .. d-2, P = prefix over 0 .. d-2
#pragma unroll
    for (int k = 0; k < FPL; ++k) P[k] = x[0][k];
#pragma unroll
    for (int j = 1; j < DMAX - 1; ++j) {
      if (j < d - 1) {
#pragma unroll
        for (int k = 0; k < FPL; ++k) { pre[j][k] = P[k]; P[k] = box_sum<ALG>(tab, P[k], x[j][k]); }
      }
    }
    // out_{d-1} = P
#pragma unroll
    for (int j = DMAX - 1; j >= 1; --j) {
      if (j == d - 1) {
        float y[FPL];
#pragma unroll
        for (int k = 0; k < FPL; ++k) y[k] = P[k];
        store(j, y);
      }
    }
    float S[FPL];
#pragma unroll
    for (int j = DMAX - 1; j >= 1; --j) {
      if (j == d - 1) {
#pragma unroll
        for (int k = 0; k < FPL; ++k) S[k] = x[j][k];
      }
    }
#pragma unroll
    for (int j = DMAX - 2; j >= 1; --j) {
      if (j <= d - 2) {
        float y[FPL];
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
          y[k] = box_sum<ALG>(tab, pre[j][k], S[k]);
          S[k] = box_sum<ALG>(tab, S[k], x[j][k]);
        }
        store(j, y);
      }
    }
    {
      float y[FPL];
#pragma unroll
      for (int k = 0; k < FPL; ++k) y[k] = S[k];
      store(0, y);
    }
  }
}

template <class F, int VMAX, int VPW, bool NT>
__global__ __launch_bounds__(kBlock) void vn_kernel(DevCode c, DevState s, DevOut out, int fused_syn, int init_pass) {
  // Tiles in the opposite order to the check kernel: workgroups are dispatched in grid order,
  // so this pass starts on the tiles whose messages the check pass wrote last (still in the
  // 256 MB Infinity Cache), and the next check pass starts where this one ended.
  const int t = s.tiles - 1 - (int)blockIdx.y;
  const int v0 = uniform((blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)) * VPW);
  if (v0 >= c.N) return;
  vn_vars<F, VMAX, VPW, NT>(c, s, out, fused_syn, init_pass, t, v0, threadIdx.x & 63);
}

template <int FPL, int DMAX>
__global__ __launch_bounds__(kBlock) void syndrome_kernel(DevCode c, DevState s) {
  __shared__ uint64_t part[kWavesPerBlock][FPL];
  syndrome_rows<FPL, DMAX>(c, s, blockIdx.y, blockIdx.x * kBlock, part);
}

template <int FPL>
__global__ __launch_bounds__(kWave) void control_kernel(DevState s, int iterations, int fused_syn, int step) {
  (void)control_tile<FPL>(s, iterations, fused_syn, step, blockIdx.x, threadIdx.x);
}

// ------------------------------------------------------------------------------------------
// Batch set-up: LLR [F][N] (row stride N) -> L tiles, optional a-priori on the first K bits
// (FeCl: parity_[i] += syst[i] in the LLR type), per-frame phase/iteration.
template <typename In, int TW>
__global__ __launch_bounds__(kBlock) void ingest_kernel(DevCode c, DevState s, const In* llr,
                                                        const In* syst, int64_t frames,
                                                        int64_t frame0, uint8_t init_phase, int fpl) {
  __shared__ float tile[64][65];
  const int t = blockIdx.y;
  const int v0 = blockIdx.x * 64;
  for (int fc = 0; fc < TW; fc += 64) {
    for (int i = threadIdx.x; i < 64 * 64; i += kBlock) {
      const int fl = i / 64, vl = i % 64;                 // read: frame-major, coalesced in v
      const int64_t f = (int64_t)t * TW + fc + fl;
      const int v = v0 + vl;
      float x = 0.0f;
      if (f < frames && v < c.N) {
        x = (float)llr[(frame0 + f) * c.N + v];
        if (syst && v < c.K) x = (float)((double)x + (double)syst[(frame0 + f) * c.K + v]);
      }
      tile[fl][vl] = x;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * 64; i += kBlock) {
      const int vl = i / 64, fl = i % 64;                 // write: variable-major, coalesced in f
      const int v = v0 + vl;
      if (v < c.N) s.L[((int64_t)t * c.N + v) * TW + fc + fl] = tile[fl][vl];
    }
    __syncthreads();
  }
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < TW; i += kBlock) {
      const int64_t f = (int64_t)t * TW + i;
      s.phase[f] = f < frames ? init_phase : (uint8_t)PH_FREE;
      s.iter[f] = 0;
    }
    if (threadIdx.x < fpl) s.fail[t * fpl + threadIdx.x] = 0;
    if (threadIdx.x == 0) {
      const int64_t live = frames - (int64_t)t * TW;
      uint32_t fl = 0;
      if (live > 0) {
        const uint8_t base = init_phase & PH_MASK;
        fl = (base == PH_RUN || base == PH_FINAL ? TF_LIVE : 0u) | (base == PH_RUN ? TF_RUN : 0u) |
             ((init_phase & PH_VNINIT) ? TF_VNINIT : 0u);
        // slots 0 .. n-1 sit in lanes 0 .. ceil(n / fpl) - 1
        const int64_t n = live < TW ? live : TW;
        const int groups = (int)((n + (int64_t)fpl * kLaneGroup - 1) / ((int64_t)fpl * kLaneGroup));
        fl |= ((1u << groups) - 1u) << TF_GROUP_SHIFT;
      }
      s.tile_flags[t] = fl;
    }
  }
}

// State [F][E] (check-to-variable, FeCl's edge order) -> message rows, rounded to the format.
// Element p of a (tile, edge) row sits in word p / PW, half p % PW.
template <typename In, class F>
__global__ __launch_bounds__(kBlock) void state_in_kernel(DevCode c, DevState s, const In* state,
                                                          int64_t frames, int64_t frame0) {
  constexpr int TW = F::TW, RW = F::RW, PW = F::PW;
  __shared__ float tile[64][65];
  const int t = blockIdx.y;
  const int e0 = blockIdx.x * 64;
  for (int fc = 0; fc < TW; fc += 64) {
    for (int i = threadIdx.x; i < 64 * 64; i += kBlock) {
      const int fl = i / 64, el = i % 64;
      const int64_t f = (int64_t)t * TW + fc + fl;
      const int e = e0 + el;
      tile[fl][el] = (f < frames && e < c.E) ? (float)state[(frame0 + f) * c.E + e] : 0.0f;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * (64 / PW); i += kBlock) {
      const int el = i / (64 / PW), wl = i % (64 / PW);
      const int e = e0 + el;
      float x[PW];
#pragma unroll
      for (int h = 0; h < PW; ++h) x[h] = tile[wl * PW + h][el];
      if (e < c.E) s.msg[((int64_t)t * c.E + e) * RW + fc / PW + wl] = F::Elem::pack(x);
    }
    __syncthreads();
  }
}

// Results of a region of tiles go to the frames its slots hold: slot j holds frame map[j]
// (-1: none), or frame j when map is null (the batch's first region).

// Launches whose grid.y counts frames or slots split into launches of at most this many rows
// (the device's reported grid limit in y is 65536).
constexpr int64_t kMaxGridY = 65535;
constexpr int64_t kUnpackBlocks = 8192;

// Hard decisions [tile][K][FPL] -> msg [F][K] bytes (0/1), and with `iters` the iteration
// count; slot j0 + blockIdx.y, bits blockIdx.x * kBlock + threadIdx.x in steps of
// gridDim.x * kBlock.  A slot whose frame is still decoding (RUN or FINAL) is skipped: it moves
// into the next region and is handed over by a later call, after its final update (every frame
// is done by the last call of a decode).
template <int FPL>
__global__ __launch_bounds__(kBlock) void unpack_hard_kernel(const uint64_t* hard, const uint8_t* phase,
                                                             const int32_t* iter, int K, int64_t frames, uint8_t* msg,
                                                             int32_t* iters, int64_t frame0, const int32_t* map,
                                                             const CompactCtl* ctl, int64_t j0) {
  constexpr int TW = kWave * FPL;
  if (ctl && !ctl->go) return;
  const int64_t j = j0 + blockIdx.y;
  const int64_t f = map ? (int64_t)map[j] : j;
  if (f < 0 || f >= frames) return;
  const int base = phase[j] & PH_MASK;
  if (base == PH_RUN || base == PH_FINAL) return;
  if (iters && blockIdx.x == 0 && threadIdx.x == 0) iters[frame0 + f] = iter[j];
  const int64_t t = j / TW;
  const int p = (int)(j % TW);
  const uint64_t* hw = hard + t * K * FPL + (p % FPL);
  uint8_t* out = msg + (frame0 + f) * K;
#pragma unroll 4
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < K; v += gridDim.x * kBlock)
    out[v] = (uint8_t)((hw[(int64_t)v * FPL] >> (p / FPL)) & 1ull);
}

// [tile][X][TW] float -> [F][count] (Out = float or double), x < count, through LDS.
template <typename Out, int TW>
__global__ __launch_bounds__(kBlock) void tiles_to_frames_kernel(const float* in, int64_t X, int64_t count,
                                                                 int64_t frames, Out* out, int64_t frame0,
                                                                 const int32_t* map, const CompactCtl* ctl) {
  __shared__ float tile[64][65];
  if (ctl && !ctl->go) return;
  const int t = blockIdx.y;
  const int64_t x0 = (int64_t)blockIdx.x * 64;
  for (int fc = 0; fc < TW; fc += 64) {
    for (int i = threadIdx.x; i < 64 * 64; i += kBlock) {
      const int xl = i / 64, fl = i % 64;
      const int64_t x = x0 + xl;
      tile[xl][fl] = x < X ? in[((int64_t)t * X + x) * TW + fc + fl] : 0.0f;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * 64; i += kBlock) {
      const int fl = i / 64, xl = i % 64;
      const int64_t j = (int64_t)t * TW + fc + fl;
      const int64_t f = map ? (int64_t)map[j] : j;
      const int64_t x = x0 + xl;
      if (f >= 0 && f < frames && x < count) out[(frame0 + f) * count + x] = (Out)tile[xl][fl];
    }
    __syncthreads();
  }
}

// Message rows -> state [F][E] (Out = float or double).
template <typename Out, class F>
__global__ __launch_bounds__(kBlock) void state_out_kernel(DevCode c, DevState s, int64_t frames,
                                                           Out* out, int64_t frame0, const int32_t* map,
                                                           const CompactCtl* ctl) {
  constexpr int TW = F::TW, RW = F::RW, PW = F::PW;
  __shared__ float tile[64][65];
  if (ctl && !ctl->go) return;
  const int t = blockIdx.y;
  const int e0 = blockIdx.x * 64;
  for (int fc = 0; fc < TW; fc += 64) {
    for (int i = threadIdx.x; i < 64 * (64 / PW); i += kBlock) {
      const int el = i / (64 / PW), wl = i % (64 / PW);
      const int e = e0 + el;
      float x[PW];
      if (e < c.E) F::Elem::unpack(s.msg[((int64_t)t * c.E + e) * RW + fc / PW + wl], x);
      else {
#pragma unroll
        for (int h = 0; h < PW; ++h) x[h] = 0.0f;
      }
#pragma unroll
      for (int h = 0; h < PW; ++h) tile[wl * PW + h][el] = x[h];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * 64; i += kBlock) {
      const int fl = i / 64, el = i % 64;
      const int64_t j = (int64_t)t * TW + fc + fl;
      const int64_t f = map ? (int64_t)map[j] : j;
      const int e = e0 + el;
      if (f >= 0 && f < frames && e < c.E) out[(frame0 + f) * c.E + e] = (Out)tile[fl][el];
    }
    __syncthreads();
  }
}

// Serving: frames admitted into free slots.  stage holds their LLRs frame-major [n][N];
// slot i of the list receives frame i (L rows written at the slot's position in its tile).
// A block moves 64 frames x 64 variables through LDS: reads run along the variables of a
// frame, writes along the slots of a variable, which are adjacent when the slots are.
template <int TW>
__global__ __launch_bounds__(kBlock) void admit_kernel(DevCode c, DevState s, const float* stage,
                                                       const int32_t* slots, int n, uint8_t init_phase) {
  __shared__ float tile[64][65];
  const int v0 = blockIdx.x * 64, f0 = blockIdx.y * 64;
  for (int i = threadIdx.x; i < 64 * 64; i += kBlock) {
    const int fl = i / 64, vl = i % 64;
    const int f = f0 + fl, v = v0 + vl;
    tile[fl][vl] = (f < n && v < c.N) ? stage[(int64_t)f * c.N + v] : 0.0f;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 64 * 64; i += kBlock) {
    const int vl = i / 64, fl = i % 64;
    const int f = f0 + fl, v = v0 + vl;
    if (f < n && v < c.N) {
      const int p = slots[f];
      s.L[((int64_t)(p / TW) * c.N + v) * TW + p % TW] = tile[fl][vl];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < 64 && f0 + (int)threadIdx.x < n) {
    const int p = slots[f0 + threadIdx.x];
    const int t = p / TW;
    s.phase[p] = init_phase;
    s.iter[p] = 0;
    const uint8_t base = init_phase & PH_MASK;
    const int group = (p % TW) / (TW / kWave) / kLaneGroup;   // slot p sits in lane (p % TW) / FPL
    atomicAnd(&s.tile_flags[t], ~TF_UNIFORM);
    atomicOr(&s.tile_flags[t], (base == PH_RUN || base == PH_FINAL ? TF_LIVE : 0u) | (base == PH_RUN ? TF_RUN : 0u) |
                                   (1u << (TF_GROUP_SHIFT + group)));
  }
}

// Serving: results of finished slots -> bits [n][K] (0/1) and iterations [n].
template <int FPL>
__global__ __launch_bounds__(kBlock) void extract_kernel(const uint64_t* hard, const int32_t* iter, int K,
                                                         const int32_t* slots, int n, uint8_t* bits,
                                                         int32_t* iters) {
  constexpr int TW = kWave * FPL;
  const int i = blockIdx.y;
  if (i >= n) return;
  const int p = slots[i];
  const int64_t t = p / TW;
  const int pos = p % TW;
  for (int v = blockIdx.x * kBlock + threadIdx.x; v < K; v += gridDim.x * kBlock) {
    const uint64_t w = hard[(t * K + v) * FPL + (pos % FPL)];
    bits[(int64_t)i * K + v] = (uint8_t)((w >> (pos / FPL)) & 1ull);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) iters[i] = iter[p];
}

// Per-frame iteration counts -> iterations [F]; n slots (map) or frames.
__global__ void iters_out_kernel(const int32_t* iter, int64_t n, int64_t frames, int32_t* out, int64_t frame0,
                                 const int32_t* map, const CompactCtl* ctl) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || (ctl && !ctl->go)) return;
  const int64_t f = map ? (int64_t)map[j] : j;
  if (f >= 0 && f < frames) out[frame0 + f] = iter[j];
}

// ------------------------------------------------------------------------------------------
// Compaction: frames that converge leave their lanes idle, and a tile works until its last
// frame is done, so late in a decode most tiles carry a few live frames.  The live frames of a
// region of tiles move, in slot order, into the first slots of a smaller region; later steps
// run on that region only.  Frames are independent, so every result is unchanged.

// Plan (one workgroup): the live slots (RUN or FINAL) of the source region become slots
// 0 .. n-1 of the destination; src_of[j] is slot j's source slot and map_out[j] its frame
// (map_in[slot], or the slot itself when map_in is null).  Destination slots n .. dst_slots-1
// get -1.  No write goes past dst_slots.
// With ctl (a captured graph, where no host sees the counts): the kernel decides.  It moves
// the frames only when they fit in fewer tiles than hold live frames now, and the tile-steps
// freed over the `remaining` steps outweigh moving each live frame once (counted as 1.5
// frame-iterations, as the eager decode counts it).  Otherwise map_out[j] = map_in[j] for
// every slot and src_of is left alone.  ctl->go and ctl->live record the outcome.
constexpr int kPlanThreads = 1024;
__global__ __launch_bounds__(kPlanThreads) void compact_plan_kernel(const uint8_t* phase, const int32_t* map_in,
                                                                    int32_t slots, int32_t* src_of, int32_t* map_out,
                                                                    int32_t dst_slots, CompactCtl* ctl, int32_t tw,
                                                                    int32_t remaining) {
  __shared__ int32_t sums[kPlanThreads];
  __shared__ int32_t tiles_live;
  const int tid = threadIdx.x;
  if (ctl) {
    if (tid == 0) tiles_live = 0;
    __syncthreads();
    int mine = 0;
    for (int t = tid; t * tw < slots; t += kPlanThreads) {
      bool any = false;
      for (int i = t * tw; i < t * tw + tw && i < slots; ++i) {
        const int b = phase[i] & PH_MASK;
        any |= b == PH_RUN || b == PH_FINAL;
      }
      mine += any ? 1 : 0;
    }
    if (mine) atomicAdd(&tiles_live, mine);
  }
  const int per = (slots + kPlanThreads - 1) / kPlanThreads;
  const int i0 = min(slots, tid * per), i1 = min(slots, i0 + per);
  int n = 0;
  for (int i = i0; i < i1; ++i) {
    const int b = phase[i] & PH_MASK;
    n += (b == PH_RUN || b == PH_FINAL) ? 1 : 0;
  }
  sums[tid] = n;
  __syncthreads();
  for (int off = 1; off < kPlanThreads; off <<= 1) {   // inclusive scan of the counts
    const int add = tid >= off ? sums[tid - off] : 0;
    __syncthreads();
    sums[tid] += add;
    __syncthreads();
  }
  const int total = sums[kPlanThreads - 1];
  if (ctl) {
    const int64_t need = (total + tw - 1) / tw, freed = (int64_t)tiles_live - need;
    const bool go = freed > 0 && 2 * freed * remaining * tw > 3 * (int64_t)total;
    if (tid == 0) {
      ctl->go = go ? 1 : 0;
      ctl->live = total;
    }
    if (!go) {
      for (int k = tid; k < slots; k += kPlanThreads) map_out[k] = map_in ? map_in[k] : k;
      return;
    }
  }
  int j = sums[tid] - n;
  for (int i = i0; i < i1; ++i) {
    const int b = phase[i] & PH_MASK;
    if (b == PH_RUN || b == PH_FINAL) {
      if (j < dst_slots) {
        src_of[j] = i;
        map_out[j] = map_in ? map_in[i] : i;
      }
      ++j;
    }
  }
  for (int k = total + tid; k < dst_slots; k += kPlanThreads) {
    src_of[k] = -1;
    map_out[k] = -1;
  }
}

template <class F>
__global__ __launch_bounds__(kBlock) void compact_msg_kernel(DevCode c, const uint32_t* src, uint32_t* dst,
                                                             const int32_t* src_of) {
  compact_msg_block<F>(c, src, dst, src_of, blockIdx.y, blockIdx.x * kCompactEdges);
}

template <int TW>
__global__ __launch_bounds__(kBlock) void compact_L_kernel(DevCode c, const float* src, float* dst,
                                                           const int32_t* src_of) {
  compact_L_block<TW>(c, src, dst, src_of, blockIdx.y, blockIdx.x * kCompactEdges);
}

__global__ void compact_slots_kernel(const uint8_t* src_phase, const int32_t* src_iter, uint8_t* dst_phase,
                                     int32_t* dst_iter, const int32_t* src_of, int32_t dst_slots) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= dst_slots) return;
  compact_slot(src_phase, src_iter, dst_phase, dst_iter, src_of, j);
}

template <int FPL>
__global__ __launch_bounds__(kWave) void retile_kernel(DevState s) {
  (void)retile_tile<FPL>(s, blockIdx.x, threadIdx.x);
}

// In place, in one launch (captured graphs).  Destination slot j < ctl->live takes source slot
// src_of[j], and src_of[j] >= j because live slots keep their order, so every source of
// destination tile tb lies in tile tb or later.  A workgroup owns a set of rows (message edges
// or channel variables) in every tile and walks the destination tiles upward: it gathers a
// tile's values into registers, meets a barrier, then stores them.  Its stores before tile tb
// went to tiles below tb, where no source of tb's slots lies, so each value is read before any
// store reaches it.  Workgroups own disjoint rows.  The last workgroup moves the phases and
// iteration counts the same way, kBlock slots at a time, then sets the tile flags.
constexpr int kInPlaceElems = 4;   // values a thread carries per destination tile

template <class F>
__global__ __launch_bounds__(kBlock) void compact_in_place_kernel(DevCode c, DevState s, const int32_t* src_of,
                                                                  const CompactCtl* ctl, int32_t msg_blocks,
                                                                  int32_t L_blocks) {
  constexpr int TW = F::TW, RW = F::RW, PW = F::PW, FPL = F::FPL;
  constexpr int kEdges = kInPlaceElems * kBlock / RW, kVars = kInPlaceElems * kBlock / TW;
  if (!ctl->go) return;
  const int64_t live = ctl->live;
  const int dtiles = (int)((live + TW - 1) / TW);
  const int b = blockIdx.x;
  if (b < msg_blocks) {
    const int e0 = b * kEdges;
    for (int tb = 0; tb < dtiles; ++tb) {
      uint32_t word[kInPlaceElems];
#pragma unroll
      for (int k = 0; k < kInPlaceElems; ++k) {
        const int i = threadIdx.x + k * kBlock, e = e0 + i / RW, w = i % RW;
        word[k] = 0;
        if (e >= c.E) continue;
#pragma unroll
        for (int h = 0; h < PW; ++h) {
          const int64_t j = (int64_t)tb * TW + w * PW + h;
          if (j >= live) continue;
          const int from = src_of[j];
          const uint32_t sw = s.msg[((int64_t)(from / TW) * c.E + e) * RW + (from % TW) / PW];
          word[k] |= (PW == 1 ? sw : (sw >> (16 * (from % PW))) & 0xffffu) << (16 * h);
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kInPlaceElems; ++k) {
        const int i = threadIdx.x + k * kBlock, e = e0 + i / RW, w = i % RW;
        if (e < c.E) s.msg[((int64_t)tb * c.E + e) * RW + w] = word[k];
      }
    }
  } else if (b < msg_blocks + L_blocks) {
    const int v0 = (b - msg_blocks) * kVars;
    for (int tb = 0; tb < dtiles; ++tb) {
      float x[kInPlaceElems];
#pragma unroll
      for (int k = 0; k < kInPlaceElems; ++k) {
        const int i = threadIdx.x + k * kBlock, v = v0 + i / TW, p = i % TW;
        const int64_t j = (int64_t)tb * TW + p;
        x[k] = 0.0f;
        if (v < c.N && j < live) {
          const int from = src_of[j];
          x[k] = s.L[((int64_t)(from / TW) * c.N + v) * TW + from % TW];
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < kInPlaceElems; ++k) {
        const int i = threadIdx.x + k * kBlock, v = v0 + i / TW, p = i % TW;
        if (v < c.N) s.L[((int64_t)tb * c.N + v) * TW + p] = x[k];
      }
    }
  } else {
    const int64_t slots = (int64_t)s.tiles * TW;
    for (int64_t j0 = 0; j0 < slots; j0 += kBlock) {
      const int64_t j = j0 + threadIdx.x;
      uint8_t ph = PH_FREE;
      int32_t it = 0;
      if (j < live) {
        const int from = src_of[j];
        ph = s.phase[from];
        it = s.iter[from];
      }
      __syncthreads();
      if (j < slots) {
        s.phase[j] = ph;
        s.iter[j] = it;
      }
    }
    __syncthreads();
    for (int t = threadIdx.x / kWave; t < s.tiles; t += kBlock / kWave) (void)retile_tile<FPL>(s, t, threadIdx.x % kWave);
  }
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Several byte ranges copied in one launch (workgroup y copies range y, grid-stride in x).
__global__ __launch_bounds__(kBlock) void copy_bytes_kernel(ByteCopies c) {
  const ByteCopy& r = c.item[blockIdx.y];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < r.n; i += (int64_t)gridDim.x * kBlock)
    r.dst[i] = r.src[i];
}

// ------------------------------------------------------------------------------------------
// Permutation::permute / dePermute of FP32 blocks (src/Permutation.h): a grid-stride loop over
// (frame, output index); dePermute zeroes its output first.
__global__ __launch_bounds__(kBlock) void permute_kernel(const int64_t* seq, int64_t n_seq, int64_t n_in,
                                                        const float* in, float* out, int64_t frames, int inverse) {
  const int64_t total = frames * n_seq;
  for (int64_t x = (int64_t)blockIdx.x * kBlock + threadIdx.x; x < total; x += (int64_t)gridDim.x * kBlock) {
    const int64_t f = x / n_seq, i = x % n_seq;
    if (inverse) out[f * n_in + seq[i]] = in[f * n_seq + i];
    else out[f * n_seq + i] = in[f * n_in + seq[i]];
  }
}

__global__ __launch_bounds__(kBlock) void zero_kernel(float* p, int64_t n) {
  for (int64_t x = (int64_t)blockIdx.x * kBlock + threadIdx.x; x < n; x += (int64_t)gridDim.x * kBlock) p[x] = 0.0f;
}

hipError_t kernels_permute(const int64_t* seq, int64_t out_size, int64_t in_size, const float* in, float* out,
                           int64_t frames, bool inverse, hipStream_t st) {
  // inverse: in is [frames][out_size] (the permuted blocks), out is [frames][in_size]
  if (inverse) {
    const int64_t n = frames * in_size;
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 65536)), dim3(kBlock), 0,
                       st, out, n);
  }
  const int64_t total = frames * out_size;
  if (total == 0) return hipGetLastError();
  hipLaunchKernelGGL(permute_kernel, dim3((unsigned)std::min<int64_t>((total + kBlock - 1) / kBlock, 65536)),
                     dim3(kBlock), 0, st, seq, out_size, in_size, in, out, frames, inverse ? 1 : 0);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Systematic encoder for staircase codes (H = [A | dual diagonal], DVB-S2): codeword =
// [msg | p] with p_r = p_{r-1} xor (xor of the message bits in row r), the accumulator form
// of FeCl's generator transform for these codes (src/detail/Ldpc.cpp:180-209).  One workgroup
// per frame: the frame's bits sit in LDS (N bytes), each thread folds a run of consecutive
// rows, a workgroup xor-scan of the runs' parities gives every run its starting p, and a
// second pass writes the parity bits; input and output move coalesced.
__global__ __launch_bounds__(kBlock) void encode_kernel(DevCode c, const uint8_t* msg, uint8_t* cw) {
  extern __shared__ uint8_t bits[];   // [0, K) message, [K, N) parity
  __shared__ uint32_t carry[kWavesPerBlock];
  const int64_t f = blockIdx.x;
  const uint8_t* m = msg + f * c.K;
  for (int i = threadIdx.x; i < c.K; i += kBlock) bits[i] = m[i] & 1;
  __syncthreads();
  const int run = (c.M + kBlock - 1) / kBlock;
  const int r0 = threadIdx.x * run, r1 = min(c.M, r0 + run);
  auto row_parity = [&](int r) {
    uint32_t x = 0;
    for (int e = c.row_ptr[r]; e < c.row_ptr[r + 1]; ++e) {
      const int v = c.col[e];
      if (v < c.K) x ^= bits[v];
    }
    return x;
  };
  uint32_t tot = 0;
  for (int r = r0; r < r1; ++r) tot ^= row_parity(r);
  // exclusive xor-scan of the runs: within the wave by shuffles, then across the waves
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint32_t inc = tot;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(inc, o, 64);
    if (lane >= o) inc ^= y;
  }
  if (lane == 63) carry[wv] = inc;
  __syncthreads();
  uint32_t x = inc ^ tot;   // exclusive within the wave
  for (int w = 0; w < wv; ++w) x ^= carry[w];
  for (int r = r0; r < r1; ++r) {
    x ^= row_parity(r);
    bits[c.K + r] = (uint8_t)x;
  }
  __syncthreads();
  uint8_t* out = cw + f * c.N;
  for (int i = threadIdx.x; i < c.N; i += kBlock) out[i] = bits[i];
}

hipError_t kernels_encode(const DevCode& c, const uint8_t* msg, uint8_t* cw, int64_t frames, hipStream_t st) {
  if (frames <= 0) return hipSuccess;
  if (c.N > 65536) return hipErrorInvalidValue;   // one frame's bits in LDS
  hipLaunchKernelGGL(encode_kernel, dim3((unsigned)frames), dim3(kBlock), (size_t)c.N, st, c, msg, cw);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Launchers (host side).  Register-staged kernels are instantiated only where the staged
// values fit the register budget (degree bound x frames per lane <= 64); the host picks a
// layout that satisfies it (layout_fits()).

template <class F, int VPW, bool NT>
static hipError_t launch_vn_w(const DevCode& c, const DevState& s, const DevOut& o, int max_dv, int fused, int init,
                              hipStream_t st) {
  constexpr int FPL = F::FPL, W = F::W;
  const int per_block = kWavesPerBlock * VPW;
  dim3 grid((c.N + per_block - 1) / per_block, s.tiles);
  if (c.vstride != var_table_stride(max_dv)) return hipErrorInvalidValue;
  // several variables per wave only where their packed rows fit (VPW * VMAX * W <= 32 words)
  if (max_dv <= 4) {
    if constexpr (VPW == 1 || VPW * 4 * W <= 32)
      hipLaunchKernelGGL((vn_kernel<F, 4, VPW, NT>), grid, dim3(kBlock), 0, st, c, s, o, fused, init);
    else return hipErrorInvalidValue;
  } else if (max_dv <= 8) {
    if constexpr (VPW == 1 || VPW * 8 * W <= 32)
      hipLaunchKernelGGL((vn_kernel<F, 8, VPW, NT>), grid, dim3(kBlock), 0, st, c, s, o, fused, init);
    else return hipErrorInvalidValue;
  } else if (max_dv <= 16) {
    if constexpr (16 * FPL <= 64 && (VPW == 1 || VPW * 16 * W <= 32))
      hipLaunchKernelGGL((vn_kernel<F, 16, VPW, NT>), grid, dim3(kBlock), 0, st, c, s, o, fused, init);
    else return hipErrorInvalidValue;
  } else if (max_dv <= 32) {
    if constexpr (32 * FPL <= 64 && VPW == 1)
      hipLaunchKernelGGL((vn_kernel<F, 32, VPW, NT>), grid, dim3(kBlock), 0, st, c, s, o, fused, init);
    else return hipErrorInvalidValue;
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class F>
static hipError_t launch_vn(const DevCode& c, const DevState& s, const DevOut& o, int max_dv, int fused, int init,
                            int vpw, bool nt, hipStream_t st) {
  if (vpw == 2) return nt ? launch_vn_w<F, 2, true>(c, s, o, max_dv, fused, init, st)
                          : launch_vn_w<F, 2, false>(c, s, o, max_dv, fused, init, st);
  if (vpw == 1) return nt ? launch_vn_w<F, 1, true>(c, s, o, max_dv, fused, init, st)
                          : launch_vn_w<F, 1, false>(c, s, o, max_dv, fused, init, st);
  return hipErrorInvalidValue;
}

template <class F, int ALG>
static hipError_t launch_cn_box(const DevCode& c, const DevState& s, const DevOpts& o, int max_dc, hipStream_t st) {
  constexpr int FPL = F::FPL;
  dim3 grid((c.M + kWavesPerBlock - 1) / kWavesPerBlock, s.tiles);
  if (max_dc <= 8) {
    if constexpr (8 * FPL <= 64) hipLaunchKernelGGL((cn_box_kernel<F, 8, ALG>), grid, dim3(kBlock), 0, st, c, s, o);
    else return hipErrorInvalidValue;
  } else if (max_dc <= 16) {
    if constexpr (16 * FPL <= 64) hipLaunchKernelGGL((cn_box_kernel<F, 16, ALG>), grid, dim3(kBlock), 0, st, c, s, o);
    else return hipErrorInvalidValue;
  } else if (max_dc <= 32) {
    if constexpr (32 * FPL <= 64) hipLaunchKernelGGL((cn_box_kernel<F, 32, ALG>), grid, dim3(kBlock), 0, st, c, s, o);
    else return hipErrorInvalidValue;
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Raw-encoding min-sum where its registers fit; false if no form fits this layout.
template <class F, bool NT>
static bool launch_cn_raw(const DevCode& c, const DevState& s, const DevOpts& o, int max_dc, hipStream_t st) {
  constexpr int W = F::W;
  dim3 grid((c.M + kWavesPerBlock - 1) / kWavesPerBlock, s.tiles);
  if (max_dc <= 8) {
    hipLaunchKernelGGL((cn_minsum_raw_kernel<F, 8, NT>), grid, dim3(kBlock), 0, st, c, s, o);
    return true;
  } else if (max_dc <= 16) {
    if constexpr (16 * W <= 64) {
      hipLaunchKernelGGL((cn_minsum_raw_kernel<F, 16, NT>), grid, dim3(kBlock), 0, st, c, s, o);
      return true;
    }
  } else if (max_dc <= 32) {
    if constexpr (32 * W <= 64) {
      hipLaunchKernelGGL((cn_minsum_raw_kernel<F, 32, NT>), grid, dim3(kBlock), 0, st, c, s, o);
      return true;
    }
  }
  return false;
}

template <class F>
static hipError_t launch_cn(const DevCode& c, const DevState& s, const DevOpts& o, int max_dc, hipStream_t st) {
  constexpr int FPL = F::FPL;
  dim3 grid((c.M + kWavesPerBlock - 1) / kWavesPerBlock, s.tiles);
  if (o.algorithm == 2) {
    bool done = false;
    if (o.sf_mode == 0) {   // raw encodings (no scaling factor to apply)
      done = (o.nt_rows & 1) ? launch_cn_raw<F, true>(c, s, o, max_dc, st) : launch_cn_raw<F, false>(c, s, o, max_dc, st);
      if (done) return hipGetLastError();
    }
    if (max_dc <= 8) {
      hipLaunchKernelGGL((cn_minsum_kernel<F, 8>), grid, dim3(kBlock), 0, st, c, s, o);
      done = true;
    } else if (max_dc <= 16) {
      if constexpr (16 * FPL <= 64) {
        hipLaunchKernelGGL((cn_minsum_kernel<F, 16>), grid, dim3(kBlock), 0, st, c, s, o);
        done = true;
      }
    } else if (max_dc <= 32) {
      if constexpr (32 * FPL <= 64) {
        hipLaunchKernelGGL((cn_minsum_kernel<F, 32>), grid, dim3(kBlock), 0, st, c, s, o);
        done = true;
      }
    }
    if (!done) {
      if (max_dc > 64) return hipErrorInvalidValue;
      hipLaunchKernelGGL((cn_minsum_kernel<F, 0>), grid, dim3(kBlock), 0, st, c, s, o);
    }
    return hipGetLastError();
  }
  if (o.algorithm == 1) return launch_cn_box<F, 1>(c, s, o, max_dc, st);
  return launch_cn_box<F, 2>(c, s, o, max_dc, st);
}

// Syndrome (unless fused into the check kernel) and the per-frame state machine.
template <int FPL>
static hipError_t launch_syn_ctrl(const DevCode& c, const DevState& s, int iterations, int fused, int step,
                                  int max_dc, hipStream_t st) {
  if (!fused) {
    dim3 grid((c.M + kBlock - 1) / kBlock, s.tiles);
    if (max_dc <= 8) hipLaunchKernelGGL((syndrome_kernel<FPL, 8>), grid, dim3(kBlock), 0, st, c, s);
    else if (max_dc <= 16) hipLaunchKernelGGL((syndrome_kernel<FPL, 16>), grid, dim3(kBlock), 0, st, c, s);
    else if (max_dc <= 32) hipLaunchKernelGGL((syndrome_kernel<FPL, 32>), grid, dim3(kBlock), 0, st, c, s);
    else hipLaunchKernelGGL((syndrome_kernel<FPL, 0>), grid, dim3(kBlock), 0, st, c, s);
  }
  hipLaunchKernelGGL((control_kernel<FPL>), dim3(s.tiles), dim3(kWave), 0, st, s, iterations, fused, step);
  return hipGetLastError();
}

template <class F>
static hipError_t launch_ingest(const DevCode& c, const DevState& s, const void* llr, const void* syst,
                                bool dbl, int64_t frames, int64_t frame0, uint8_t init_phase, hipStream_t st) {
  dim3 grid((c.N + 63) / 64, s.tiles);
  if (dbl)
    hipLaunchKernelGGL((ingest_kernel<double, F::TW>), grid, dim3(kBlock), 0, st, c, s, (const double*)llr,
                       (const double*)syst, frames, frame0, init_phase, F::FPL);
  else
    hipLaunchKernelGGL((ingest_kernel<float, F::TW>), grid, dim3(kBlock), 0, st, c, s, (const float*)llr,
                       (const float*)syst, frames, frame0, init_phase, F::FPL);
  return hipGetLastError();
}

template <class F>
static hipError_t launch_state_in(const DevCode& c, const DevState& s, const void* state, bool dbl,
                                  int64_t frames, int64_t frame0, hipStream_t st) {
  dim3 grid((c.E + 63) / 64, s.tiles);
  if (dbl)
    hipLaunchKernelGGL((state_in_kernel<double, F>), grid, dim3(kBlock), 0, st, c, s, (const double*)state, frames, frame0);
  else
    hipLaunchKernelGGL((state_in_kernel<float, F>), grid, dim3(kBlock), 0, st, c, s, (const float*)state, frames, frame0);
  return hipGetLastError();
}

template <class F>
static hipError_t launch_results(const DevCode& c, const DevState& s, const DevOut& out, int64_t frames,
                                 int64_t frame0, const BatchResults& r, bool dbl, const int32_t* map,
                                 const CompactCtl* ctl, hipStream_t st) {
  constexpr int FPL = F::FPL, TW = F::TW;
  const int64_t n = map ? (int64_t)s.tiles * TW : frames;   // slots of the region, or frames
  if (n <= 0) return hipSuccess;
  if (r.msg_bits) {
    // about kUnpackBlocks workgroups in all (at least one per slot), each striding over the bits:
    // a hand-over that finds few finished frames, or none, dispatches few empty workgroups
    const int64_t kb = (c.K + kBlock - 1) / kBlock;
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>(kb, kUnpackBlocks / n));
    for (int64_t j0 = 0; j0 < n; j0 += kMaxGridY)
      hipLaunchKernelGGL((unpack_hard_kernel<FPL>), dim3(gx, (unsigned)std::min(kMaxGridY, n - j0)), dim3(kBlock), 0, st,
                         out.hard, s.phase, s.iter, c.K, frames, r.msg_bits, r.iterations, frame0, map, ctl, j0);
  } else if (r.iterations) {
    hipLaunchKernelGGL(iters_out_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s.iter, n, frames,
                       r.iterations, frame0, map, ctl);
  }
  auto to_frames = [&](const float* in, int64_t X, int64_t count, void* dst) {
    dim3 grid((unsigned)((count + 63) / 64), s.tiles);
    if (dbl) hipLaunchKernelGGL((tiles_to_frames_kernel<double, TW>), grid, dim3(kBlock), 0, st, in, X, count,
                                frames, (double*)dst, frame0, map, ctl);
    else hipLaunchKernelGGL((tiles_to_frames_kernel<float, TW>), grid, dim3(kBlock), 0, st, in, X, count,
                            frames, (float*)dst, frame0, map, ctl);
  };
  if (r.msg) to_frames(out.app, c.K, c.K, r.msg);
  if (r.parity) to_frames(out.ext, c.N, c.N, r.parity);
  if (r.syst) to_frames(out.ext, c.N, c.K, r.syst);
  if (r.state) {
    if (dbl) hipLaunchKernelGGL((state_out_kernel<double, F>), dim3((c.E + 63) / 64, s.tiles), dim3(kBlock), 0, st,
                                c, s, frames, (double*)r.state, frame0, map, ctl);
    else hipLaunchKernelGGL((state_out_kernel<float, F>), dim3((c.E + 63) / 64, s.tiles), dim3(kBlock), 0, st,
                            c, s, frames, (float*)r.state, frame0, map, ctl);
  }
  return hipGetLastError();
}

// Moves the frames of slots src_of[j] of `from` into slots j of `to` (src_of[j] = -1: empty).
template <class F>
static hipError_t launch_compact(const DevCode& c, const DevState& from, const DevState& to, const int32_t* map_in,
                                 int32_t* map_out, int32_t* src_of, hipStream_t st) {
  constexpr int FPL = F::FPL, TW = F::TW;
  const int32_t slots = from.tiles * TW, dst_slots = to.tiles * TW;
  hipLaunchKernelGGL(compact_plan_kernel, dim3(1), dim3(kPlanThreads), 0, st, from.phase, map_in, slots, src_of,
                     map_out, dst_slots, (CompactCtl*)nullptr, TW, 0);
  hipLaunchKernelGGL((compact_msg_kernel<F>), dim3((c.E + kCompactEdges - 1) / kCompactEdges, to.tiles), dim3(kBlock),
                     0, st, c, from.msg, to.msg, (const int32_t*)src_of);
  hipLaunchKernelGGL((compact_L_kernel<TW>), dim3((c.N + kCompactEdges - 1) / kCompactEdges, to.tiles), dim3(kBlock),
                     0, st, c, from.L, to.L, (const int32_t*)src_of);
  hipLaunchKernelGGL(compact_slots_kernel, dim3((dst_slots + 255) / 256), dim3(256), 0, st, from.phase, from.iter,
                     to.phase, to.iter, (const int32_t*)src_of, dst_slots);
  hipLaunchKernelGGL((retile_kernel<FPL>), dim3(to.tiles), dim3(kWave), 0, st, to);
  return hipGetLastError();
}

template <class F>
static hipError_t launch_compact_in_place(const DevCode& c, const DevState& s, const int32_t* src_of,
                                          const CompactCtl* ctl, hipStream_t st) {
  constexpr int kEdges = kInPlaceElems * kBlock / F::RW, kVars = kInPlaceElems * kBlock / F::TW;
  static_assert(kEdges >= 1 && kVars >= 1, "a workgroup owns at least one row");
  const int32_t msg_blocks = (c.E + kEdges - 1) / kEdges, L_blocks = (c.N + kVars - 1) / kVars;
  hipLaunchKernelGGL((compact_in_place_kernel<F>), dim3(msg_blocks + L_blocks + 1), dim3(kBlock), 0, st, c, s, src_of,
                     ctl, msg_blocks, L_blocks);
  return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Format / layout dispatch.
hipError_t kernels_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t kernels_copy_bytes(const ByteCopies& c, hipStream_t st) {
  if (c.count <= 0) return hipSuccess;
  if (c.count > kMaxByteCopies) return hipErrorInvalidValue;
  int64_t longest = 0;
  for (int i = 0; i < c.count; ++i) longest = std::max<int64_t>(longest, c.item[i].n);
  const unsigned blocks = (unsigned)std::min<int64_t>(64, (longest + kBlock - 1) / kBlock);
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(copy_bytes_kernel, dim3(blocks, (unsigned)c.count), dim3(kBlock), 0, st, c);
  return hipGetLastError();
}

#define FECL_DISPATCH_W(E, lay, CALL)                  \
  switch (lay.words) {                                 \
    case 1: { using F = Fmt<E, 1>; return CALL; }      \
    case 2: { using F = Fmt<E, 2>; return CALL; }      \
    case 4: { using F = Fmt<E, 4>; return CALL; }      \
    default: return hipErrorInvalidValue;              \
  }
#define FECL_DISPATCH(lay, CALL)                                    \
  switch (lay.fmt) {                                                \
    case 0: FECL_DISPATCH_W(ElemF32, lay, CALL)                     \
    case 1: FECL_DISPATCH_W(ElemF16, lay, CALL)                     \
    case 2: FECL_DISPATCH_W(ElemBF16, lay, CALL)                    \
    default: return hipErrorInvalidValue;                           \
  }

static int degree_bound(int d) { return d <= 4 ? 4 : d <= 8 ? 8 : d <= 16 ? 16 : d <= 32 ? 32 : 1 << 20; }

bool layout_fits(const Layout& lay, int algorithm, int max_dc, int max_dv) {
  const int fpl = lay.fpl();
  if (max_dv > 32 || (max_dv > 8 && degree_bound(max_dv) * fpl > 64)) return false;
  if (algorithm == 2) return max_dc <= 64;  // the streaming form covers rows above the budget
  return max_dc <= 32 && degree_bound(max_dc) * fpl <= 64;
}

bool cn_register_staged(const Layout& lay, int algorithm, int sf_mode, int max_dc) {
  const int db = degree_bound(max_dc);
  if (algorithm == 2 && sf_mode == 0) return db * lay.words <= 64;   // raw-encoding min-sum
  return db * lay.fpl() <= 64;
}

bool vpw_fits(const Layout& lay, int max_dv) {
  if (lay.vpw == 1) return true;
  return lay.vpw == 2 && max_dv <= 16 && 2 * degree_bound(max_dv) * lay.words <= 32;
}

hipError_t kernels_ingest(const Layout& lay, const DevCode& c, const DevState& s, const void* llr, const void* syst,
                          bool dbl, int64_t frames, int64_t frame0, uint8_t init_phase, hipStream_t st) {
  FECL_DISPATCH(lay, launch_ingest<F>(c, s, llr, syst, dbl, frames, frame0, init_phase, st));
}
hipError_t kernels_state_in(const Layout& lay, const DevCode& c, const DevState& s, const void* state, bool dbl,
                            int64_t frames, int64_t frame0, hipStream_t st) {
  FECL_DISPATCH(lay, launch_state_in<F>(c, s, state, dbl, frames, frame0, st));
}
hipError_t kernels_vn_init(const Layout& lay, const DevCode& c, const DevState& s, int max_dv, hipStream_t st) {
  FECL_DISPATCH(lay, launch_vn<F>(c, s, DevOut{}, max_dv, 0, 1, lay.vpw, false, st));
}
hipError_t kernels_check(const Layout& lay, const DevCode& c, const DevState& s, const DevOpts& o, int max_dc,
                         hipStream_t st) {
  FECL_DISPATCH(lay, launch_cn<F>(c, s, o, max_dc, st));
}
hipError_t kernels_variable(const Layout& lay, const DevCode& c, const DevState& s, const DevOut& out, int max_dv,
                            int fused, bool nt, hipStream_t st) {
  FECL_DISPATCH(lay, launch_vn<F>(c, s, out, max_dv, fused, 0, lay.vpw, nt, st));
}
hipError_t kernels_syndrome_control(const Layout& lay, const DevCode& c, const DevState& s, int iterations,
                                    int fused, int step, int max_dc, hipStream_t st) {
  switch (lay.fpl()) {
    case 1: return launch_syn_ctrl<1>(c, s, iterations, fused, step, max_dc, st);
    case 2: return launch_syn_ctrl<2>(c, s, iterations, fused, step, max_dc, st);
    case 4: return launch_syn_ctrl<4>(c, s, iterations, fused, step, max_dc, st);
    case 8: return launch_syn_ctrl<8>(c, s, iterations, fused, step, max_dc, st);
    default: return hipErrorInvalidValue;
  }
}
hipError_t kernels_admit(const Layout& lay, const DevCode& c, const DevState& s, const float* stage,
                         const int32_t* slots, int n, uint8_t init_phase, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  dim3 grid((unsigned)((c.N + 63) / 64), (unsigned)((n + 63) / 64));
  switch (lay.tw()) {
    case 64: hipLaunchKernelGGL((admit_kernel<64>), grid, dim3(kBlock), 0, st, c, s, stage, slots, n, init_phase); break;
    case 128: hipLaunchKernelGGL((admit_kernel<128>), grid, dim3(kBlock), 0, st, c, s, stage, slots, n, init_phase); break;
    case 256: hipLaunchKernelGGL((admit_kernel<256>), grid, dim3(kBlock), 0, st, c, s, stage, slots, n, init_phase); break;
    case 512: hipLaunchKernelGGL((admit_kernel<512>), grid, dim3(kBlock), 0, st, c, s, stage, slots, n, init_phase); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t kernels_extract(const Layout& lay, const DevCode& c, const DevState& s, const DevOut& out,
                           const int32_t* slots, int n, uint8_t* bits, int32_t* iters, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (lay.fpl() != 1 && lay.fpl() != 2 && lay.fpl() != 4 && lay.fpl() != 8) return hipErrorInvalidValue;
  for (int i0 = 0; i0 < n; i0 += (int)kMaxGridY) {   // slots i0 .. of the list, at most kMaxGridY a launch
    const int m = (int)std::min<int64_t>(kMaxGridY, n - i0);
    const int32_t* sl = slots + i0;
    uint8_t* b = bits + (int64_t)i0 * c.K;
    int32_t* it = iters + i0;
    dim3 grid((unsigned)std::min<int64_t>((c.K + kBlock - 1) / kBlock, 64), (unsigned)m);
    switch (lay.fpl()) {
      case 1: hipLaunchKernelGGL((extract_kernel<1>), grid, dim3(kBlock), 0, st, out.hard, s.iter, c.K, sl, m, b, it); break;
      case 2: hipLaunchKernelGGL((extract_kernel<2>), grid, dim3(kBlock), 0, st, out.hard, s.iter, c.K, sl, m, b, it); break;
      case 4: hipLaunchKernelGGL((extract_kernel<4>), grid, dim3(kBlock), 0, st, out.hard, s.iter, c.K, sl, m, b, it); break;
      default: hipLaunchKernelGGL((extract_kernel<8>), grid, dim3(kBlock), 0, st, out.hard, s.iter, c.K, sl, m, b, it); break;
    }
  }
  return hipGetLastError();
}

hipError_t kernels_results(const Layout& lay, const DevCode& c, const DevState& s, const DevOut& out, int64_t frames,
                           int64_t frame0, const BatchResults& r, bool dbl, const int32_t* map, const CompactCtl* ctl,
                           hipStream_t st) {
  FECL_DISPATCH(lay, launch_results<F>(c, s, out, frames, frame0, r, dbl, map, ctl, st));
}

hipError_t kernels_compact(const Layout& lay, const DevCode& c, const DevState& from, const DevState& to,
                           const int32_t* map_in, int32_t* map_out, int32_t* src_of, hipStream_t st) {
  if (from.tiles <= 0 || to.tiles <= 0 || (int64_t)from.tiles * lay.tw() > INT32_MAX) return hipErrorInvalidValue;
  FECL_DISPATCH(lay, launch_compact<F>(c, from, to, map_in, map_out, src_of, st));
}

hipError_t kernels_compact_plan(const Layout& lay, const DevState& s, const int32_t* map_in, int32_t* map_out,
                                int32_t* src_of, CompactCtl* ctl, int32_t remaining, hipStream_t st) {
  const int64_t slots = (int64_t)s.tiles * lay.tw();
  if (s.tiles <= 0 || slots > INT32_MAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(compact_plan_kernel, dim3(1), dim3(kPlanThreads), 0, st, s.phase, map_in, (int32_t)slots, src_of,
                     map_out, (int32_t)slots, ctl, (int32_t)lay.tw(), remaining);
  return hipGetLastError();
}

hipError_t kernels_compact_in_place(const Layout& lay, const DevCode& c, const DevState& s, const int32_t* src_of,
                                    const CompactCtl* ctl, hipStream_t st) {
  if (s.tiles <= 0 || !ctl || !src_of) return hipErrorInvalidValue;
  FECL_DISPATCH(lay, launch_compact_in_place<F>(c, s, src_of, ctl, st));
}

}  // namespace fecl
