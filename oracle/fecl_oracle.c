/* CPU oracle (test infrastructure only; see fecl_oracle.h for what pins each function). */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fecl_oracle.h"
#include "../fecl_amd/csrc/dvbs2_tables.h"

/* ---- DVB-S2 construction (ETSI EN 302 307-1 sec. 5.3.2; reference: src/DvbS2.cpp:1595-1645,
 * whose index bookkeeping is right only for n=64800 rate 1/2).  With q from the table,
 * M = 360*q checks and K = n - M information bits: information bit i of group g = i/360
 * joins check (a + (i mod 360)*q) mod M for every address a in table row g; parity bit K+j
 * joins checks j and j+1 (the last one only check M-1).  Rows list their columns in
 * ascending variable order, as SparseBitMatrix::transpose() leaves them. */
int orc_dvbs2_build(int n, int rate_num, int rate_den, int64_t* dims, int32_t* row_ptr, int32_t* col) {
  const fecl_dvbs2_code_entry* ce = NULL;
  for (int i = 0; i < 21; ++i)
    if (fecl_dvbs2_codes[i].n == n && fecl_dvbs2_codes[i].rate_num == rate_num &&
        fecl_dvbs2_codes[i].rate_den == rate_den)
      ce = &fecl_dvbs2_codes[i];
  if (!ce) return -1;
  const int64_t q = ce->q, M = 360 * q, K = n - M;
  /* count edges per check */
  int64_t* deg = (int64_t*)calloc((size_t)M, sizeof(int64_t));
  const uint16_t* rec = fecl_dvbs2_addr + ce->offset;
  for (int g = 0; g < ce->table_rows; ++g) {
    int cnt = rec[0];
    for (int k = 0; k < 360; ++k)
      for (int j = 0; j < cnt; ++j) deg[(rec[1 + j] + k * q) % M]++;
    rec += 1 + cnt;
  }
  for (int64_t j = 0; j < M; ++j) { deg[j]++; if (j + 1 < M) deg[j + 1]++; }
  int64_t E = 0;
  for (int64_t r = 0; r < M; ++r) E += deg[r];
  if (dims) { dims[0] = M; dims[1] = n; dims[2] = K; dims[3] = E; }
  if (row_ptr && col) {
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)M);
    row_ptr[0] = 0;
    for (int64_t r = 0; r < M; ++r) { row_ptr[r + 1] = (int32_t)(row_ptr[r] + deg[r]); fill[r] = row_ptr[r]; }
    /* variables in ascending order: information bits, then parity bits */
    rec = fecl_dvbs2_addr + ce->offset;
    for (int g = 0; g < ce->table_rows; ++g) {
      int cnt = rec[0];
      for (int k = 0; k < 360; ++k) {
        int64_t v = (int64_t)g * 360 + k;
        for (int j = 0; j < cnt; ++j) col[fill[(rec[1 + j] + k * q) % M]++] = (int32_t)v;
      }
      rec += 1 + cnt;
    }
    for (int64_t j = 0; j < M; ++j) {
      col[fill[j]++] = (int32_t)(K + j);
      if (j + 1 < M) col[fill[j + 1]++] = (int32_t)(K + j);
    }
    free(fill);
  }
  free(deg);
  return 0;
}

/* Staircase encoder: p_r = p_{r-1} xor (xor of the information bits on check r). */
int orc_dvbs2_encode(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N,
                     const uint8_t* info, uint8_t* cw, int64_t frames) {
  const int64_t K = N - M;
  for (int64_t f = 0; f < frames; ++f) {
    const uint8_t* u = info + f * K;
    uint8_t* c = cw + f * N;
    memcpy(c, u, (size_t)K);
    uint8_t prev = 0;
    for (int64_t r = 0; r < M; ++r) {
      uint8_t s = 0;
      for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e)
        if (col[e] < K) s ^= u[col[e]] & 1;
      prev ^= s;
      c[K + r] = prev;
    }
  }
  return 0;
}

int64_t orc_syndrome_weight(const int32_t* row_ptr, const int32_t* col, int64_t M, const uint8_t* cw) {
  int64_t w = 0;
  for (int64_t r = 0; r < M; ++r) {
    uint8_t s = 0;
    for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) s ^= cw[col[e]] & 1;
    w += s;
  }
  return w;
}

/* ---- 16-bit message storage: saturate to the largest finite value, then round to nearest
 * even.  Written from the IEEE-754 binary16 and bfloat16 layouts. */
float orc_quantize(float x, int fmt) {
  if (fmt == ORC_MSG_F32 || isnan(x)) return x;
  uint32_t u;
  if (fmt == ORC_MSG_F16) {
    if (x > 65504.0f) x = 65504.0f;
    if (x < -65504.0f) x = -65504.0f;
    float ax = fabsf(x);
    if (ax < 6.103515625e-05f) {          /* below 2^-14: binary16 subnormal, quantum 2^-24 */
      float r = nearbyintf(ax * 16777216.0f) / 16777216.0f;
      return signbit(x) ? -r : r;
    }
    memcpy(&u, &x, 4);
    u += 0xfffu + ((u >> 13) & 1u);       /* keep 10 of 23 mantissa bits */
    u &= ~0x1fffu;
  } else {
    const float bmax = 3.38953138925153547590e38f;  /* 0x7f7f0000 */
    if (x > bmax) x = bmax;
    if (x < -bmax) x = -bmax;
    memcpy(&u, &x, 4);
    u += 0x7fffu + ((u >> 16) & 1u);      /* keep 7 of 23 mantissa bits */
    u &= 0xffff0000u;
  }
  memcpy(&x, &u, 4);
  return x;
}

/* ---- the decoder, once per LLR type ---- */
#define REAL double
#define SFX _f64
#include "bp_body.h"
#undef REAL
#undef SFX
#define REAL float
#define SFX _f32
#include "bp_body.h"
#undef REAL
#undef SFX

int orc_bp_decode(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, int64_t K,
                  const double* llr, int64_t frames, int alg, int iterations,
                  const double* sf, int sf_rows, int sf_cols, int prec, int msg_fmt,
                  const double* syst_in, const double* state_in,
                  uint8_t* hard, double* msg, double* parity, double* syst, double* state,
                  int32_t* iters, int nthreads) {
  return orc_bp_decode_diag(row_ptr, col, M, N, K, llr, frames, alg, iterations, sf, sf_rows, sf_cols, prec,
                            msg_fmt, syst_in, state_in, hard, msg, parity, syst, state, iters, NULL, nthreads);
}

int orc_bp_decode_diag(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, int64_t K,
                       const double* llr, int64_t frames, int alg, int iterations,
                       const double* sf, int sf_rows, int sf_cols, int prec, int msg_fmt,
                       const double* syst_in, const double* state_in,
                       uint8_t* hard, double* msg, double* parity, double* syst, double* state,
                       int32_t* iters, uint64_t* lin_side, int nthreads) {
  if (iterations < 1 || frames < 0 || sf_rows < 1 || alg < 0 || alg > 3) return -1;
  if (prec != 32 && prec != 64) return -1;
  if (prec == 64 && msg_fmt != ORC_MSG_F32) return -1;
  for (int64_t r = 0; r < M; ++r) {
    int64_t d = row_ptr[r + 1] - row_ptr[r];
    if (d < 2 || d - 1 > sf_cols) return -1;
  }
  if (frames == 0) return 0;
  const int64_t E = row_ptr[M];
  if (prec == 64) {
    code_t_f64 cd = {row_ptr, col, M, N, K, E, alg, iterations, msg_fmt, sf, sf_rows, sf_cols};
    return decode_f64(&cd, llr, frames, syst_in, state_in, hard, msg, parity, syst, state, iters, lin_side,
                      nthreads);
  }
  code_t_f32 cd = {row_ptr, col, M, N, K, E, alg, iterations, msg_fmt, sf, sf_rows, sf_cols};
  return decode_f32(&cd, llr, frames, syst_in, state_in, hard, msg, parity, syst, state, iters, lin_side,
                    nthreads);
}
