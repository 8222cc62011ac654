// Device-side building blocks shared by the per-step kernels (ldpc_kernels.hip) and the
// persistent multi-code kernel (ldpc_persist.hip): message formats, row and variable updates,
// syndrome and per-frame control, each as a function of its (tile, row / variable) so that
// both launch forms run the same arithmetic.
//
// Reference arithmetic (FeCl):
//   check update     src/detail/BpDecoder/BpDecoderImpl.cpp:133-159
//   bit update       src/detail/BpDecoder/BpDecoderImpl.cpp:162-179
//   schedule/output  src/detail/BpDecoder/BpDecoderImpl.cpp:38-130
//   MinBoxSum / LinearBoxSum / BoxSum   src/detail/LlrMetrics.h:126-196
//   linear table     src/detail/LinearTable.h:57-81
//   syndrome test    src/detail/Ldpc.cpp:160-172
#pragma once
#include <hip/hip_runtime.h>

#include "ldpc_kernels.h"

namespace fecl {

// ------------------------------------------------------------------------------------------
// Element formats: PW messages per 32-bit word.
// 16-bit storage: saturate to the format's largest finite value (v_med3_f32), then round to
// nearest even with gfx950's two-lane converts (v_cvt_pk_f16_f32 / v_cvt_pk_bf16_f32).
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

constexpr float kF16Max = 65504.0f;
constexpr float kBF16Max = 3.38953138925153547590e38f;   // 0x1.fep127

__device__ static inline uint32_t f32x2_to_f16x2(float a, float b) {
  f32x2 x = {__builtin_amdgcn_fmed3f(a, -kF16Max, kF16Max), __builtin_amdgcn_fmed3f(b, -kF16Max, kF16Max)};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(x, f16x2));
}
__device__ static inline uint32_t f32x2_to_bf16x2(float a, float b) {
  f32x2 x = {__builtin_amdgcn_fmed3f(a, -kBF16Max, kBF16Max), __builtin_amdgcn_fmed3f(b, -kBF16Max, kBF16Max)};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(x, bf16x2));
}
__device__ static inline float f16_bits_to_f32(uint32_t b) {
  return (float)__builtin_bit_cast(_Float16, (uint16_t)b);
}
struct ElemF32 {
  static constexpr int PW = 1;
  __device__ static inline void unpack(uint32_t w, float* x) { x[0] = __uint_as_float(w); }
  __device__ static inline uint32_t pack(const float* x) { return __float_as_uint(x[0]); }
  __device__ static inline float quantize(float x) { return x; }
};
struct ElemF16 {
  static constexpr int PW = 2;
  __device__ static inline void unpack(uint32_t w, float* x) {
    x[0] = f16_bits_to_f32(w & 0xffffu);
    x[1] = f16_bits_to_f32(w >> 16);
  }
  __device__ static inline uint32_t pack(const float* x) { return f32x2_to_f16x2(x[0], x[1]); }
  __device__ static inline float quantize(float x) { return f16_bits_to_f32(f32x2_to_f16x2(x, 0.0f)); }
};
struct ElemBF16 {
  static constexpr int PW = 2;
  __device__ static inline void unpack(uint32_t w, float* x) {
    x[0] = __uint_as_float(w << 16);
    x[1] = __uint_as_float(w & 0xffff0000u);
  }
  __device__ static inline uint32_t pack(const float* x) { return f32x2_to_bf16x2(x[0], x[1]); }
  __device__ static inline float quantize(float x) { return __uint_as_float(f32x2_to_bf16x2(x, 0.0f) << 16); }
};

// Message rows are read once and written once per kernel.  NT = true marks those accesses
// non-temporal, which pays off when a batch's rows far exceed the caches; smaller batches keep
// the default policy, under which the rows one kernel writes are partly still in the 256 MB
// Infinity Cache when the next kernel reads them.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
template <int W, bool NT = false>
__device__ static inline void load_words(const uint32_t* p, uint32_t (&w)[W]) {
  if constexpr (W == 1) {
    w[0] = NT ? __builtin_nontemporal_load(p) : p[0];
  } else if constexpr (W == 2) {
    const u32x2* q = reinterpret_cast<const u32x2*>(p);
    const u32x2 v = NT ? __builtin_nontemporal_load(q) : *q;
    w[0] = v.x; w[1] = v.y;
  } else {
    static_assert(W == 4, "W is 1, 2 or 4");
    const u32x4* q = reinterpret_cast<const u32x4*>(p);
    const u32x4 v = NT ? __builtin_nontemporal_load(q) : *q;
    w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
  }
}
template <int W, bool NT = false>
__device__ static inline void store_words(uint32_t* p, const uint32_t (&w)[W]) {
  if constexpr (W == 1) {
    if constexpr (NT) __builtin_nontemporal_store(w[0], p);
    else p[0] = w[0];
  } else if constexpr (W == 2) {
    const u32x2 v = {w[0], w[1]};
    if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<u32x2*>(p));
    else *reinterpret_cast<u32x2*>(p) = v;
  } else {
    const u32x4 v = {w[0], w[1], w[2], w[3]};
    if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(p));
    else *reinterpret_cast<u32x4*>(p) = v;
  }
}
template <int N>
__device__ static inline void load_floats(const float* p, float (&x)[N]) {
  if constexpr (N == 1) {
    x[0] = p[0];
  } else if constexpr (N == 2) {
    const float2 v = *reinterpret_cast<const float2*>(p);
    x[0] = v.x; x[1] = v.y;
  } else {
#pragma unroll
    for (int i = 0; i < N; i += 4) {
      const float4 v = *reinterpret_cast<const float4*>(p + i);
      x[i] = v.x; x[i + 1] = v.y; x[i + 2] = v.z; x[i + 3] = v.w;
    }
  }
}

template <class E, int W_>
struct Fmt {
  using Elem = E;
  static constexpr int W = W_, PW = E::PW, FPL = W_ * E::PW, TW = kWave * FPL, RW = kWave * W_;
  template <bool NT = false>
  __device__ static inline void load(const uint32_t* p, float (&x)[FPL]) {
    uint32_t w[W];
    load_words<W, NT>(p, w);
#pragma unroll
    for (int i = 0; i < W; ++i) E::unpack(w[i], &x[i * PW]);
  }
  template <bool NT = false>
  __device__ static inline void store(uint32_t* p, const float (&x)[FPL]) {
    uint32_t w[W];
#pragma unroll
    for (int i = 0; i < W; ++i) w[i] = E::pack(&x[i * PW]);
    store_words<W, NT>(p, w);
  }
  __device__ static inline float quantize(float x) { return E::quantize(x); }
};

__device__ static inline int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ static inline uint64_t wave_or(uint64_t x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x |= __shfl_xor(x, o, 64);
  return x;
}

// Scaling factor of check update t for a lane (FeCl: sf * post(...) with a double sf,
// rounded to the message type).  mode 0: 1.0; 1: float product equals the rounded double one.
__device__ static inline float apply_sf(const DevOpts& o, int row, int dcol, float v) {
  if (o.sf_mode == 0) return v;
  if (o.sf_mode == 1) return o.sf32[row * o.sf_cols + dcol] * v;
  return (float)(o.sf[row * o.sf_cols + dcol] * (double)v);
}

// ------------------------------------------------------------------------------------------
// Lane groups of a tile (TF_GROUP_SHIFT): the 64-bit mask of lanes holding a live frame folded
// to one bit per kLaneGroup lanes.
__device__ static inline uint32_t lane_groups(uint64_t lanes) {
  static_assert(kLaneGroup == 4 && TF_GROUP_SHIFT + kWave / kLaneGroup <= 32, "16 groups of 4 lanes");
  lanes |= lanes >> 1;
  lanes |= lanes >> 2;   // bit 4g: any of lanes 4g .. 4g+3
  uint32_t g = 0;
#pragma unroll
  for (int i = 0; i < kWave / kLaneGroup; ++i) g |= (uint32_t)((lanes >> (kLaneGroup * i)) & 1u) << i;
  return g;
}
// The lane whose words this lane's row loads fetch: itself if its group may hold a live frame,
// else the first lane of the first such group, whose lines the wave reads anyway.  A lane of
// another group holds no frame that takes an update, so it stores nothing, and what it
// computes from those words is never kept; its loads then add no memory traffic.  This is
// where tiles that hold few frames (small batches, the last frames of a batch after a
// compaction, a serving pool's partly filled tiles) save their row traffic.
__device__ static inline int lane_for_loads(uint32_t tflags, int lane) {
  const uint32_t g = tflags >> TF_GROUP_SHIFT;
  const int first = kLaneGroup * __builtin_ctz(g | (1u << (kWave / kLaneGroup - 1)));
  return ((g >> (lane / kLaneGroup)) & 1u) ? lane : first;
}

// Per-lane view of the frames a check-node wavefront serves.
template <int FPL>
struct CnLanes {
  bool act[FPL];     // frame takes a check update this step (RUN or FINAL)
  bool fresh[FPL];   // its v2c is still L[col(e)]
  int sfrow[FPL];    // scaling-factor row: iteration count, or I-1 for the final update
  bool lane_any, lane_all, wave_fresh, wave_seen;  // wave_seen: some frame has a prior VN
};

// Frame state loads and the decisions drawn from them are split so that a kernel can issue
// its row loads in between: the wait for the state then does not wait for the rows.
// A TF_UNIFORM tile (every frame RUN, no pending flag, one iteration count) skips the
// per-frame loads.
template <int FPL>
__device__ static inline void lane_state(const DevState& s, int t, uint32_t tflags, int lane, uint8_t (&ph)[FPL],
                                         int (&it)[FPL]) {
  constexpr int TW = kWave * FPL;
  if (tflags & TF_UNIFORM) {
    const int ti = uniform(s.tile_iter[t]);
#pragma unroll
    for (int k = 0; k < FPL; ++k) { ph[k] = PH_RUN; it[k] = ti; }
    return;
  }
#pragma unroll
  for (int k = 0; k < FPL; ++k) {
    const int64_t p = (int64_t)t * TW + lane * FPL + k;
    ph[k] = s.phase[p];
    it[k] = s.iter[p];
  }
}

template <int FPL>
__device__ static inline bool cn_decide(const DevOpts& o, const uint8_t (&ph)[FPL], const int (&it)[FPL],
                                        CnLanes<FPL>& L) {
  L.lane_any = false;
  L.lane_all = true;
  bool any_fresh = false, seen = false;
#pragma unroll
  for (int k = 0; k < FPL; ++k) {
    const int base = ph[k] & PH_MASK;
    L.act[k] = base == PH_RUN || base == PH_FINAL;
    L.fresh[k] = L.act[k] && (ph[k] & PH_FRESH);
    L.sfrow[k] = (base == PH_FINAL ? o.iterations - 1 : it[k]) % o.sf_rows;
    L.lane_any |= L.act[k];
    L.lane_all &= L.act[k];
    any_fresh |= L.fresh[k];
    seen |= base == PH_RUN && it[k] >= 1;
  }
  if (__ballot(L.lane_any) == 0) return false;
  L.wave_fresh = __ballot(any_fresh) != 0;
  L.wave_seen = __ballot(seen) != 0;
  return true;
}

// Fused syndrome (used when the scaling factors do not change with the iteration, so the
// check update that follows a successful syndrome equals FeCl's final one): the row's parity
// over the hard decisions of the previous variable update, or-ed into the tile's fail mask.
// Only bits not yet set are sent as atomics, so a tile costs a few atomics per frame.
template <int FPL>
__device__ static inline void cn_syndrome(const DevCode& c, const DevState& s, int t, int e0, int d, int lane) {
  // lanes 0..d-1 each fetch one column's hard words; a wave-wide xor folds them
  uint64_t syn[FPL];
#pragma unroll
  for (int k = 0; k < FPL; ++k) syn[k] = 0;
  if (lane < d) {
    const int v = c.col[e0 + lane];
    const uint64_t* hv = s.hard + ((int64_t)t * c.N + v) * FPL;
#pragma unroll
    for (int k = 0; k < FPL; ++k) syn[k] = hv[k];
  }
#pragma unroll
  for (int k = 0; k < FPL; ++k) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) syn[k] ^= __shfl_xor(syn[k], o, 64);
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < FPL; ++k) {
      if (!syn[k]) continue;
      uint64_t* f = &s.fail[t * FPL + k];
      const uint64_t seen = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (syn[k] & ~seen) atomicOr((unsigned long long*)f, (unsigned long long)syn[k]);
    }
  }
}

// ------------------------------------------------------------------------------------------
// Check-node update, min-sum (DecoderAlgorithm::Approximate).  MinBoxSum is sign * min and is
// associative, so FeCl's forward/backward pass equals "min of the others, xor of the others'
// signs" exactly: output j has magnitude min_{i != j} |x_i| and sign bit
// ((d-2) & 1) ^ xor_{i != j} signbit(x_i)  (a box-sum of two operands sets the sign bit
// 1 ^ s_a ^ s_b).  The row keeps the two smallest magnitudes m1 <= m2 (branch-free:
// m2 = min(m2, max(m1, a)), m1 = min(m1, a)); "min of the others" is m2 where |x_j| == m1 and
// m1 elsewhere (a tie makes m2 == m1, so either choice is exact).  One wavefront per
// (check row, tile).  DMAX > 0: the row's d words are loaded into registers at once
// (d <= DMAX); DMAX == 0: two streaming passes (any d <= 64).
template <class F, int DMAX>
__device__ static inline void cn_minsum_row(const DevCode& c, const DevState& s, const DevOpts& o, int t, int r,
                                            int lane) {
  constexpr int FPL = F::FPL, TW = F::TW, RW = F::RW;
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
  float xs[DMAX > 0 ? DMAX : 1][FPL];
  if constexpr (DMAX > 0) {
#pragma unroll
    for (int j = 0; j < DMAX; ++j)
      if (j < d) F::load(mld + (int64_t)j * RW, xs[j]);
  }
  CnLanes<FPL> ln;
  if (!cn_decide<FPL>(o, ph, it, ln)) return;

  float m1[FPL], m2[FPL];
  uint64_t sg[FPL];
#pragma unroll
  for (int k = 0; k < FPL; ++k) { m1[k] = __builtin_inff(); m2[k] = __builtin_inff(); sg[k] = 0; }
  auto accumulate = [&](int j, const float (&x)[FPL]) {
#pragma unroll
    for (int k = 0; k < FPL; ++k) {
      const float a = fabsf(x[k]);
      sg[k] |= (uint64_t)(__float_as_uint(x[k]) >> 31) << j;
      m2[k] = fminf(m2[k], fmaxf(m1[k], a));
      m1[k] = fminf(m1[k], a);
    }
  };
  auto fresh_x = [&](int j, float (&x)[FPL]) {   // frames whose v2c is still L[col(e)]
    if (ln.wave_fresh) {
      const int v = uniform(c.col[e0 + j]);
      float l[FPL];
      load_floats<FPL>(Lt + (int64_t)v * TW, l);
#pragma unroll
      for (int k = 0; k < FPL; ++k)
        if (ln.fresh[k]) x[k] = F::quantize(l[k]);
    }
  };
  if constexpr (DMAX > 0) {
#pragma unroll
    for (int j = 0; j < DMAX; ++j)
      if (j < d) fresh_x(j, xs[j]);
  }
  if (o.fused_syn && ln.wave_seen) cn_syndrome<FPL>(c, s, t, e0, d, lane);
  if constexpr (DMAX > 0) {
#pragma unroll
    for (int j = 0; j < DMAX; ++j)
      if (j < d) accumulate(j, xs[j]);
  } else {
    for (int j = 0; j < d; ++j) {
      float x[FPL];
      F::load(mld + (int64_t)j * RW, x);
      fresh_x(j, x);
      accumulate(j, x);
    }
  }
  float m1s[FPL], m2s[FPL];
  uint32_t par[FPL];
#pragma unroll
  for (int k = 0; k < FPL; ++k) {
    par[k] = (uint32_t)(((d - 2) & 1) ^ (__popcll(sg[k]) & 1));
    m1s[k] = apply_sf(o, ln.sfrow[k], d - 2, m1[k]);
    m2s[k] = apply_sf(o, ln.sfrow[k], d - 2, m2[k]);
  }
  if (!ln.lane_any) return;
  // x: the row's input word as accumulated (frames that take no update are never fresh, so
  // for them x is the stored word, which they keep)
  auto emit = [&](int j, const float (&x)[FPL]) {
    float y[FPL];
#pragma unroll
    for (int k = 0; k < FPL; ++k) {
      const float mag = (fabsf(x[k]) == m1[k]) ? m2s[k] : m1s[k];
      const uint32_t sb = par[k] ^ (uint32_t)((sg[k] >> j) & 1);
      y[k] = ln.act[k] ? (sb ? -mag : mag) : x[k];
    }
    F::store(mrow + (int64_t)j * RW, y);
  };
  if constexpr (DMAX > 0) {
#pragma unroll
    for (int j = 0; j < DMAX; ++j)
      if (j < d) emit(j, xs[j]);
  } else {
    for (int j = 0; j < d; ++j) {
      float x[FPL];
      F::load(mld + (int64_t)j * RW, x);
      fresh_x(j, x);
      emit(j, x);
    }
  }
}


// Min-sum without scaling factors: every output is one of the row's stored values with a new
// sign, so the update runs on the raw encodings with integer min/max (for FP16/BF16, two frames
// per 32-bit word).  For non-NaN values the magnitude order is the integer order of the bits
// without the sign bit, and min-sum never forms a NaN.  Same results as cn_minsum_kernel, bit
// for bit.
template <class F>
struct RawMinSum;
template <int W>
struct RawMinSum<Fmt<ElemF32, W>> {
  using M = uint32_t;
  static constexpr uint32_t kMag = 0x7fffffffu, kSign = 0x80000000u;
  __device__ static inline M mag(uint32_t w) { return w & kMag; }
  __device__ static inline M init() { return kMag; }
  __device__ static inline void fold(M& m1, M& m2, M a) { m2 = min(m2, max(m1, a)); m1 = min(m1, a); }
  // |x_j| == m1 -> m2, else m1
  __device__ static inline uint32_t pick(M a, M m1, M m2) { return a == m1 ? m2 : m1; }
  __device__ static inline uint32_t lanes(const bool* f) { return f[0] ? 0xffffffffu : 0u; }
};
template <class E, int W>
struct RawMinSum16 {
  using M = u16x2;
  static constexpr uint32_t kMag = 0x7fff7fffu, kSign = 0x80008000u;
  __device__ static inline M mag(uint32_t w) { return __builtin_bit_cast(u16x2, w & kMag); }
  __device__ static inline M init() { return u16x2{0x7fff, 0x7fff}; }
  __device__ static inline void fold(M& m1, M& m2, M a) {
    m2 = __builtin_elementwise_min(m2, __builtin_elementwise_max(m1, a));
    m1 = __builtin_elementwise_min(m1, a);
  }
  // a >= m1, so a - m1 does not wrap; 1 - (a - m1) saturates to 1 where a == m1, else 0
  __device__ static inline uint32_t pick(M a, M m1, M m2) {
    const u16x2 one = {1, 1};
    return __builtin_bit_cast(uint32_t, m1 + __builtin_elementwise_sub_sat(one, a - m1) * (m2 - m1));
  }
  __device__ static inline uint32_t lanes(const bool* f) { return (f[0] ? 0x0000ffffu : 0u) | (f[1] ? 0xffff0000u : 0u); }
};
template <int W>
struct RawMinSum<Fmt<ElemF16, W>> : RawMinSum16<ElemF16, W> {};
template <int W>
struct RawMinSum<Fmt<ElemBF16, W>> : RawMinSum16<ElemBF16, W> {};

template <class F, int DMAX, bool NT>
__device__ static inline void cn_minsum_raw_row(const DevCode& c, const DevState& s, const DevOpts& o, int t, int r,
                                                int lane) {
  static_assert(DMAX > 0, "register-staged rows");
  using R = RawMinSum<F>;
  using M = typename R::M;
  constexpr int FPL = F::FPL, TW = F::TW, RW = F::RW, W = F::W, PW = F::PW;
  // the row's extent and the tile's flags are requested together (the exit test reads both,
  // which keeps the compiler from deferring the extent load past it)
  const int e0 = uniform(c.row_ptr[r]);
  const int d = uniform(c.row_ptr[r + 1]) - e0;
  const uint32_t tflags = uniform((int)s.tile_flags[t]);
  if (!(tflags & TF_LIVE) || d < 0) return;
  const int ll = lane_for_loads(tflags, lane);
  uint32_t* mrow = s.msg + ((int64_t)t * c.E + e0) * RW + lane * W;   // stores
  const uint32_t* mld = mrow + (ll - lane) * W;                        // row loads
  const float* Lt = s.L + (int64_t)t * c.N * TW + ll * FPL;
  uint8_t ph[FPL];
  int it[FPL];
  lane_state<FPL>(s, t, tflags, lane, ph, it);
  uint32_t w[DMAX][W];
#pragma unroll
  for (int j = 0; j < DMAX; ++j)
    if (j < d) load_words<W, NT>(mld + (int64_t)j * RW, w[j]);
  CnLanes<FPL> ln;
  if (!cn_decide<FPL>(o, ph, it, ln)) return;
  if (ln.wave_fresh) {   // frames whose v2c is still L[col(e)], rounded to the storage format
#pragma unroll
    for (int j = 0; j < DMAX; ++j) {
      if (j < d) {
        const int v = uniform(c.col[e0 + j]);
        float l[FPL];
        load_floats<FPL>(Lt + (int64_t)v * TW, l);
#pragma unroll
        for (int i = 0; i < W; ++i) {
          const uint32_t q = F::Elem::pack(&l[PW * i]);
          const uint32_t fm = R::lanes(&ln.fresh[PW * i]);
          w[j][i] = (q & fm) | (w[j][i] & ~fm);
        }
      }
    }
  }
  if (o.fused_syn && ln.wave_seen) cn_syndrome<FPL>(c, s, t, e0, d, lane);
  M m1[W], m2[W];
  uint32_t sx[W];
#pragma unroll
  for (int i = 0; i < W; ++i) { m1[i] = R::init(); m2[i] = m1[i]; sx[i] = 0; }
#pragma unroll
  for (int j = 0; j < DMAX; ++j) {
    if (j < d) {
#pragma unroll
      for (int i = 0; i < W; ++i) {
        R::fold(m1[i], m2[i], R::mag(w[j][i]));
        sx[i] ^= w[j][i];
      }
    }
  }
  if (!ln.lane_any) return;
  const uint32_t flip = ((d - 2) & 1) ? R::kSign : 0u;
  uint32_t am[W];
#pragma unroll
  for (int i = 0; i < W; ++i) am[i] = R::lanes(&ln.act[PW * i]);
#pragma unroll
  for (int j = 0; j < DMAX; ++j) {
    if (j < d) {
      uint32_t y[W];
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const uint32_t out = R::pick(R::mag(w[j][i]), m1[i], m2[i]) | ((sx[i] ^ flip ^ w[j][i]) & R::kSign);
        y[i] = (out & am[i]) | (w[j][i] & ~am[i]);   // frames without an update keep their word
      }
      store_words<W, NT>(mrow + (int64_t)j * RW, y);
    }
  }
}


// ------------------------------------------------------------------------------------------
// Exact (BoxSum, src/detail/LlrMetrics.h:126-140): prior p = tanh(-x/2), product of the other
// inputs' priors, post -log((1+p)/(1-p)).  Evaluated on the complement q = 1 - |p| with the
// sign bits taken apart:
//   q = 1 - tanh(|x|/2) = 2u/(1+u), u = e^-|x|;   1 - q = (1-q_a)(1-q_b)  =>  q = q_a + q_b (1-q_a);
//   |out| = log((1+|p|)/(1-|p|)) = log((2-q)/q).
// This is BoxSum's value without the cancellation of 1-p as |p| -> 1 and without tanh's
// saturation, and it costs two transcendentals per input and two per output (a pairwise
// log-domain box-sum needs four each).  |x| is clamped to 87 so that u stays a normal number:
// that changes q by less than 2e-38, which moves no output unless every other input of the row
// is as large, and such an output is still at least 87 - log(d-1).  Every q stays in [0, 1],
// so (2-q)/q >= 1 up to rounding; the sign bit is cleared before the output's sign is set.
__device__ static inline float exact_prior(float x) {
  const float u = __expf(-fminf(fabsf(x), 87.0f));
  return (2.0f * u) * __builtin_amdgcn_rcpf(1.0f + u);
}
__device__ static inline float exact_comb(float a, float b) { return a + b * (1.0f - a); }
__device__ static inline float exact_post(float q, uint32_t sign) {
  const float m = __logf((2.0f - q) * __builtin_amdgcn_rcpf(q));
  return __uint_as_float((__float_as_uint(m) & 0x7fffffffu) | (sign & 0x80000000u));
}

// ------------------------------------------------------------------------------------------
// Pairwise box-sums for the Linear check update (LinearBoxSum: FeCl sign convention, the
// result has magnitude ~min(|a|,|b|) and sign -sign(a)*sign(b)): min - f(|a+b|) + f(|a-b|)
// with FeCl's 8-knot table f.
// The Linear table's knots y[q] and slopes s[q] sit in LDS (filled once per workgroup), so a
// lookup is one ds_read_b64 instead of a chain of selects.  Entry 7 is (0, 0): FeCl's f is 0
// from x = 7 on, so clamping x to 7.5 replaces the test; x - i is fract(x), exact for
// 0 <= x < 8, and the product and sum round separately as FeCl's (y[i+1]-y[i])*(x-i)+y[i].
__device__ static inline float lin_f(const float2* tab, float z) {
  const float x = fminf(z * 2.0f, 7.5f);
  const float2 ys = tab[(int)x];
  const float prod = ys.y * __builtin_amdgcn_fractf(x);
  return prod + ys.x;
}
template <int ALG>
__device__ static inline float box_sum(const float2* tab, float a, float b) {
  static_assert(ALG == 1, "pairwise box-sums serve the Linear rule only");
  const float m = fminf(fabsf(a), fabsf(b));
  const float s1 = fabsf(a + b), s2 = fabsf(a - b);
  const float f1 = lin_f(tab, s1), f2 = lin_f(tab, s2);
  // +m where the signs differ, -m where they agree (m's own sign bit is clear)
  const uint32_t same = ~(__float_as_uint(a) ^ __float_as_uint(b)) & 0x80000000u;
  const float t = __uint_as_float(__float_as_uint(m) | same) - f1;
  return t + f2;
}


// ------------------------------------------------------------------------------------------
// Variable-node update in FeCl's order (BpDecoderImpl.cpp:162-179):
//   v2c_j = (0 + c_1 + ... + c_{j-1}) + (L + c_dv + ... + c_{j+1}),  APP = L + c_dv + ... + c_1
// with edges of the variable in ascending edge order, then hard = APP >= 0 per frame.
// One wavefront per (VPW consecutive variables, tile): the tile's frame state is read once and
// every variable's rows are requested before the first is used.  The same pass forms the
// outputs of frames whose final check update ran in this step (BpDecoderImpl.cpp:61-69,
// :109-129): with the final c2v,
//   decode   : app = L + c_1 + ... + c_dv (ascending), bit = app >= 0
//   soDecode : ext = 0 + c_1 + ... + c_dv, msg = L + ext
// and leaves their messages (the final state) untouched.  init_pass builds v2c from a
// supplied state (BpDecoderImpl.cpp:81-88) for frames flagged PH_VNINIT.
template <class F, int VMAX, int VPW, bool NT>
__device__ static inline void vn_vars(const DevCode& c, const DevState& s, const DevOut& out, int fused_syn,
                                      int init_pass, int t, int v0, int lane) {
  constexpr int FPL = F::FPL, TW = F::TW, RW = F::RW, W = F::W;
  // the tile's flags and the variables' edge lists (VPW * VMAX ids; the table is padded to a
  // whole number of waves) are requested together (the exit test reads both, which keeps the
  // compiler from deferring the list load past it); channel values, frame state and rows
  // follow at once
  int eid[VPW][VMAX];
#pragma unroll
  for (int i = 0; i < VPW; ++i)
#pragma unroll
    for (int j = 0; j < VMAX; ++j) eid[i][j] = c.var_tab[(v0 + i) * VMAX + j];
  const uint32_t tflags = uniform((int)s.tile_flags[t]);
  if (!(tflags & (init_pass ? TF_VNINIT : TF_LIVE)) || eid[0][0] < -1) return;
  const int ll = lane_for_loads(tflags, lane);
  float Lv[VPW][FPL];
#pragma unroll
  for (int i = 0; i < VPW; ++i)
    if (v0 + i < c.N) load_floats<FPL>(s.L + ((int64_t)t * c.N + v0 + i) * TW + ll * FPL, Lv[i]);
  int dv[VPW];
#pragma unroll
  for (int i = 0; i < VPW; ++i) {
    dv[i] = 0;
#pragma unroll
    for (int j = 0; j < VMAX; ++j) dv[i] += eid[i][j] >= 0 ? 1 : 0;
  }
  uint8_t ph[FPL];
  int itv[FPL];
  lane_state<FPL>(s, t, tflags, lane, ph, itv);
  uint64_t fw[FPL];
#pragma unroll
  for (int k = 0; k < FPL; ++k) fw[k] = s.fail[t * FPL + k];
  uint32_t* mt = s.msg + (int64_t)t * c.E * RW + lane * W;   // stores
  const uint32_t* mtl = mt + (ll - lane) * W;                 // row loads
  uint32_t wv[VPW][VMAX][W];
#pragma unroll
  for (int i = 0; i < VPW; ++i)
#pragma unroll
    for (int j = 0; j < VMAX; ++j)
      if (j < dv[i]) load_words<W, NT>(mtl + (int64_t)eid[i][j] * RW, wv[i][j]);
  bool run[FPL], fin[FPL];
  bool lane_run = false, lane_fin = false;
#pragma unroll
  for (int k = 0; k < FPL; ++k) {
    const int base = ph[k] & PH_MASK;
    // fused syndrome: a RUN frame whose previous hard decision satisfied every check has just
    // had its final check update (see cn_syndrome)
    const bool passed = fused_syn && !init_pass && base == PH_RUN && itv[k] >= 1 && ((fw[k] >> lane) & 1ull) == 0;
    run[k] = init_pass ? ((ph[k] & PH_VNINIT) != 0) : (base == PH_RUN && !passed);
    fin[k] = !init_pass && (base == PH_FINAL || passed);
    lane_run |= run[k];
    lane_fin |= fin[k];
  }
  const bool wave_run = __ballot(lane_run) != 0;
  const bool wave_fin = __ballot(lane_fin) != 0;
  if (!wave_run && !wave_fin) return;
#pragma unroll
  for (int i = 0; i < VPW; ++i) {
    const int v = v0 + i;
    if (v >= c.N) break;
    float cv[VMAX][FPL];
#pragma unroll
    for (int j = 0; j < VMAX; ++j) {
      if (j < dv[i]) {
#pragma unroll
        for (int q = 0; q < W; ++q) F::Elem::unpack(wv[i][j][q], &cv[j][q * F::PW]);
      }
    }
    if (wave_fin) {
      float app[FPL], ext[FPL];
#pragma unroll
      for (int k = 0; k < FPL; ++k) { app[k] = Lv[i][k]; ext[k] = 0.0f; }
#pragma unroll
      for (int j = 0; j < VMAX; ++j) {
        if (j < dv[i]) {
#pragma unroll
          for (int k = 0; k < FPL; ++k) { app[k] = app[k] + cv[j][k]; ext[k] = ext[k] + cv[j][k]; }
        }
      }
      if (v < c.K) {
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
          const uint64_t bits = __ballot(app[k] >= 0.0f);
          const uint64_t fm = __ballot(fin[k]);
          if (lane == 0 && fm) {
            uint64_t* w = &out.hard[((int64_t)t * c.K + v) * FPL + k];
            *w = (*w & ~fm) | (bits & fm);
          }
        }
      }
      if (out.ext) {
#pragma unroll
        for (int k = 0; k < FPL; ++k)
          if (fin[k]) out.ext[((int64_t)t * c.N + v) * TW + lane * FPL + k] = ext[k];
      }
      if (out.app && v < c.K) {
#pragma unroll
        for (int k = 0; k < FPL; ++k)
          if (fin[k]) out.app[((int64_t)t * c.K + v) * TW + lane * FPL + k] = Lv[i][k] + ext[k];
      }
    }
    if (!wave_run) continue;
    float pre[VMAX][FPL];
    float acc[FPL];
#pragma unroll
    for (int k = 0; k < FPL; ++k) acc[k] = 0.0f;
#pragma unroll
    for (int j = 0; j < VMAX; ++j) {
      if (j < dv[i]) {
#pragma unroll
        for (int k = 0; k < FPL; ++k) { pre[j][k] = acc[k]; acc[k] = acc[k] + cv[j][k]; }
      }
    }
#pragma unroll
    for (int k = 0; k < FPL; ++k) acc[k] = Lv[i][k];
#pragma unroll
    for (int j = VMAX - 1; j >= 0; --j) {
      if (j < dv[i]) {
        float y[FPL];
#pragma unroll
        for (int k = 0; k < FPL; ++k) {
          y[k] = pre[j][k] + acc[k];
          acc[k] = acc[k] + cv[j][k];
          if (!run[k]) y[k] = cv[j][k];   // other frames keep their word (exact round trip)
        }
        if (lane_run) F::template store<NT>(mt + (int64_t)eid[i][j] * RW, y);
      }
    }
    if (init_pass) continue;
#pragma unroll
    for (int k = 0; k < FPL; ++k) {
      const uint64_t m = __ballot(acc[k] >= 0.0f);
      if (lane == 0) s.hard[((int64_t)t * c.N + v) * FPL + k] = m;
    }
  }
}


// ------------------------------------------------------------------------------------------
// Syndrome (when not fused into the check kernel): one thread per (check row, tile) xors the
// hard words of its columns (all loads of a row issued together); the failing frames are or-ed
// across the workgroup in LDS and sent as one atomic per workgroup into fail[tile].
// One workgroup's rows [r0, r0 + kBlock) of tile t (every thread of the workgroup calls it;
// part: kWavesPerBlock * FPL words of LDS).
template <int FPL, int DMAX>
__device__ static inline void syndrome_rows(const DevCode& c, const DevState& s, int t, int r0,
                                            uint64_t (*part)[FPL]) {
  if (!(s.tile_flags[t] & TF_RUN)) return;
  const int r = r0 + threadIdx.x;
  uint64_t f[FPL];
#pragma unroll
  for (int k = 0; k < FPL; ++k) f[k] = 0;
  if (r < c.M) {
    const uint64_t* ht = s.hard + (int64_t)t * c.N * FPL;
    const int e0 = c.row_ptr[r];
    const int d = c.row_ptr[r + 1] - e0;
    if constexpr (DMAX > 0) {
      int v[DMAX];
#pragma unroll
      for (int j = 0; j < DMAX; ++j)
        if (j < d) v[j] = c.col[e0 + j];
#pragma unroll
      for (int j = 0; j < DMAX; ++j) {
        if (j < d) {
#pragma unroll
          for (int k = 0; k < FPL; ++k) f[k] ^= ht[(int64_t)v[j] * FPL + k];
        }
      }
    } else {
      for (int j = 0; j < d; ++j) {
        const int v = c.col[e0 + j];
#pragma unroll
        for (int k = 0; k < FPL; ++k) f[k] ^= ht[(int64_t)v * FPL + k];
      }
    }
  }
  const int wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < FPL; ++k) {
    const uint64_t w = wave_or(f[k]);
    if ((threadIdx.x & 63) == 0) part[wave][k] = w;
  }
  __syncthreads();
  if (threadIdx.x < FPL) {
    uint64_t w = 0;
#pragma unroll
    for (int i = 0; i < kWavesPerBlock; ++i) w |= part[i][threadIdx.x];
    if (w) atomicOr((unsigned long long*)&s.fail[t * FPL + threadIdx.x], (unsigned long long)w);
  }
}


// ------------------------------------------------------------------------------------------
// Per-frame state machine after a step (one wavefront per tile).  FeCl's loop
//   for t in 0 .. I-2: { CN(t); VN; if syndrome ok: break }  CN(I-1)
// becomes: RUN frames count the update; they turn FINAL when the syndrome passes or the
// I-1 regular updates are used up; FINAL frames (their last update used factor row I-1)
// turn DONE.  With the fused syndrome, the syndrome seen in a step is the one of the previous
// variable update, so a passing RUN frame has just had its final update and turns DONE.
// The iteration count is the number of check-node updates executed.
// One wavefront (lane = 0..63) updates tile t; returns whether a frame of the tile is still
// live (RUN or FINAL), the same on every lane.
template <int FPL>
__device__ static inline bool control_tile(const DevState& s, int iterations, int fused_syn, int step, int t,
                                           int lane) {
  constexpr int TW = kWave * FPL;
  bool any_live = false, any_run = false, all_run = true;
  int alive = 0;
  int iters[FPL];
#pragma unroll
  for (int k = 0; k < FPL; ++k) {
    const uint64_t fw = s.fail[t * FPL + k];
    const int64_t p = (int64_t)t * TW + lane * FPL + k;
    uint8_t ph = s.phase[p];
    const int base = ph & PH_MASK;
    iters[k] = -1;
    if (base == PH_RUN) {
      const int it0 = s.iter[p];
      const int it = it0 + 1;
      s.iter[p] = it;
      iters[k] = it;
      const bool ok = ((fw >> lane) & 1ull) == 0;
      if (fused_syn)
        ph = (it0 >= 1 && ok) ? PH_DONE : (it >= iterations - 1 ? PH_FINAL : PH_RUN);
      else
        ph = (ok || it >= iterations - 1) ? PH_FINAL : PH_RUN;
    } else if (base == PH_FINAL) {
      s.iter[p] = s.iter[p] + 1;
      ph = PH_DONE;
    } else {
      ph = base;
    }
    s.phase[p] = ph;
    any_live |= ph == PH_RUN || ph == PH_FINAL;
    any_run |= ph == PH_RUN;
    all_run &= ph == PH_RUN;
    alive += (ph == PH_RUN || ph == PH_FINAL) ? 1 : 0;
  }
  // uniform: every frame still RUN (phases leave this kernel without pending flags) at one
  // iteration count
  const int it_ref = __builtin_amdgcn_readfirstlane(iters[0]);
#pragma unroll
  for (int k = 0; k < FPL; ++k) all_run &= iters[k] == it_ref;
  const bool uni = __ballot(!all_run) == 0;
  const uint64_t blive = __ballot(any_live), brun = __ballot(any_run);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) alive += __shfl_xor(alive, o, 64);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < FPL; ++k) s.fail[t * FPL + k] = 0;
    s.tile_flags[t] = (blive ? TF_LIVE : 0u) | (brun ? TF_RUN : 0u) | (uni ? TF_UNIFORM : 0u) |
                      (lane_groups(blive) << TF_GROUP_SHIFT);
    s.tile_iter[t] = it_ref;
    if (alive && step >= 0) atomicAdd(&s.active[step], alive);
    if (s.live_out) s.live_out[t] = alive;
  }
  return blive != 0;
}


// ------------------------------------------------------------------------------------------
// Compaction moves (see ldpc_kernels.hip for the plan).
// Message rows: destination word w of a (tile, edge) row gathers its PW elements (element p of
// a row sits in word p / PW, bits 16 * (p % PW) up) from their source slots; empty slots get 0.
// A workgroup covers kCompactEdges edges of one destination tile.
constexpr int kCompactEdges = 16;

// Edges [e_begin, e_begin + kCompactEdges) of destination tile tb (every thread of the
// workgroup calls it).
template <class F>
__device__ static inline void compact_msg_block(const DevCode& c, const uint32_t* src, uint32_t* dst,
                                                const int32_t* src_of, int tb, int e_begin) {
  constexpr int TW = F::TW, RW = F::RW, PW = F::PW;
  for (int i = threadIdx.x; i < kCompactEdges * RW; i += kBlock) {
    const int e = e_begin + i / RW, w = i % RW;
    if (e >= c.E) break;
    uint32_t word = 0;
#pragma unroll
    for (int h = 0; h < PW; ++h) {
      const int from = src_of[(int64_t)tb * TW + w * PW + h];
      if (from >= 0) {
        const int ta = from / TW, pa = from % TW;
        const uint32_t sw = src[((int64_t)ta * c.E + e) * RW + pa / PW];
        const uint32_t x = PW == 1 ? sw : (sw >> (16 * (pa % PW))) & 0xffffu;
        word |= x << (16 * h);
      }
    }
    dst[((int64_t)tb * c.E + e) * RW + w] = word;
  }
}

// Channel values [tile][N][TW]: variables [v_begin, v_begin + kCompactEdges) of destination
// tile tb.
template <int TW>
__device__ static inline void compact_L_block(const DevCode& c, const float* src, float* dst, const int32_t* src_of,
                                              int tb, int v_begin) {
  for (int i = threadIdx.x; i < kCompactEdges * TW; i += kBlock) {
    const int v = v_begin + i / TW, p = i % TW;
    if (v >= c.N) break;
    const int from = src_of[(int64_t)tb * TW + p];
    float x = 0.0f;
    if (from >= 0) x = src[((int64_t)(from / TW) * c.N + v) * TW + from % TW];
    dst[((int64_t)tb * c.N + v) * TW + p] = x;
  }
}

// Phase and iteration count of destination slot j (an empty slot: PH_FREE).
__device__ static inline void compact_slot(const uint8_t* src_phase, const int32_t* src_iter, uint8_t* dst_phase,
                                           int32_t* dst_iter, const int32_t* src_of, int j) {
  const int from = src_of[j];
  dst_phase[j] = from >= 0 ? src_phase[from] : (uint8_t)PH_FREE;
  dst_iter[j] = from >= 0 ? src_iter[from] : 0;
}

// Tile flags of tile t from its phases (one wavefront, lane = 0..63), as the control kernel
// leaves them at the end of a step; fail masks cleared.  Returns whether a frame is live.
template <int FPL>
__device__ static inline bool retile_tile(const DevState& s, int t, int lane) {
  constexpr int TW = kWave * FPL;
  bool any_live = false, any_run = false, all_run = true;
  int iters[FPL];
#pragma unroll
  for (int k = 0; k < FPL; ++k) {
    const int64_t p = (int64_t)t * TW + lane * FPL + k;
    const uint8_t ph = s.phase[p];
    const int base = ph & PH_MASK;
    iters[k] = base == PH_RUN ? s.iter[p] : -1;
    any_live |= base == PH_RUN || base == PH_FINAL;
    any_run |= base == PH_RUN;
    all_run &= ph == PH_RUN;
  }
  const int it_ref = __builtin_amdgcn_readfirstlane(iters[0]);
#pragma unroll
  for (int k = 0; k < FPL; ++k) all_run &= iters[k] == it_ref;
  const bool uni = __ballot(!all_run) == 0;
  const uint64_t blive = __ballot(any_live), brun = __ballot(any_run);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < FPL; ++k) s.fail[t * FPL + k] = 0;
    s.tile_flags[t] = (blive ? TF_LIVE : 0u) | (brun ? TF_RUN : 0u) | (uni ? TF_UNIFORM : 0u) |
                      (lane_groups(blive) << TF_GROUP_SHIFT);
    s.tile_iter[t] = it_ref;
  }
  return blive != 0;
}

}  // namespace fecl
