/* Decoder body, included twice by fecl_oracle.c with REAL = double / float and SFX the name
 * suffix.  Each step follows FeCl's operation order and type conversions exactly, since the
 * results are compared bit for bit:
 *   prior/sum/post   src/detail/LlrMetrics.h:126-196 (BoxSum, LinearBoxSum, MinBoxSum)
 *   linear table     src/detail/LinearTable.h:57-81
 *   check update     src/detail/BpDecoder/BpDecoderImpl.cpp:133-159
 *   bit update       src/detail/BpDecoder/BpDecoderImpl.cpp:162-179
 *   block schedule   src/detail/BpDecoder/BpDecoderImpl.cpp:38-130
 *   syndrome test    src/detail/Ldpc.cpp:160-172
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* Exact: prior tanh(-x/2), sum a*b, post -log((1+x)/(1-x)); the literals promote to double. */
static inline REAL FN(prior_exact)(REAL x) { return (REAL)tanh(-x / 2.0); }
static inline REAL FN(post_exact)(REAL x) { return (REAL)(-log((1.0 + x) / (1.0 - x))); }

/* y[i] = log(1 + exp(-i/2)), i = 0..7, computed in double and stored as REAL. */
static REAL FN(lin_y)[8];
static void FN(lin_init)(void) {
  for (int i = 0; i < 8; ++i) FN(lin_y)[i] = (REAL)log(1.0 + exp(-((double)(REAL)i / (double)(REAL)2.0)));
}
/* Which side of the table's cutoff (x >= 7, where f drops from y[7] = log(1 + e^-3.5) to 0)
 * every lookup of the current frame fell on, folded into a 64-bit FNV-1a hash in lookup order
 * (a diagnostic: two decodes of a frame that run the same lookups agree on every side iff
 * their hashes agree, up to hash collisions). */
static __thread uint64_t FN(lin_side);
/* Piecewise-linear log(1+exp(-x)): x *= 2; 0 beyond the last knot; interpolate. */
static inline REAL FN(lin_f)(REAL x) {
  x *= (REAL)2.0;
  FN(lin_side) = (FN(lin_side) ^ (uint64_t)(x >= (REAL)7)) * 0x100000001b3ull;
  if (x >= (REAL)7) return 0;
  size_t i = (size_t)x;
  return (FN(lin_y)[i + 1] - FN(lin_y)[i]) * (x - (REAL)i) + FN(lin_y)[i];
}
static inline REAL FN(rmin)(REAL a, REAL b) { return (b < a) ? b : a; }  /* std::min */
static inline REAL FN(sum_linear)(REAL a, REAL b) {
  REAL m = FN(rmin)((REAL)fabs(a), (REAL)fabs(b));
  REAL f1 = FN(lin_f)((REAL)fabs(a + b));
  REAL f2 = FN(lin_f)((REAL)fabs(a - b));
  if (signbit(a) ^ signbit(b)) {
    REAL t = m - f1;
    return t + f2;
  }
  REAL t = -m - f1;
  return t + f2;
}
static inline REAL FN(sum_min)(REAL a, REAL b) {
  double m = FN(rmin)((REAL)fabs(a), (REAL)fabs(b));
  if ((signbit(a) != 0) ^ (signbit(b) != 0)) return (REAL)m;
  return (REAL)(-m);
}

/* Exact box-plus evaluated in the LLR domain (not one of FeCl's algorithms: the same
 * mathematical value as BoxSum without tanh's saturation, used to measure conditioning):
 * -s_a s_b min(|a|,|b|) - log1p(e^-|a+b|) + log1p(e^-|a-b|). */
static inline REAL FN(sum_stable)(REAL a, REAL b) {
  REAL m = FN(rmin)((REAL)fabs(a), (REAL)fabs(b));
  REAL f1 = (REAL)log1p(exp(-(double)fabs(a + b)));
  REAL f2 = (REAL)log1p(exp(-(double)fabs(a - b)));
  REAL t = ((signbit(a) != 0) ^ (signbit(b) != 0)) ? m - f1 : -m - f1;
  return t + f2;
}

static inline REAL FN(prior)(int alg, REAL x) { return alg == ORC_EXACT ? FN(prior_exact)(x) : x; }
static inline REAL FN(post)(int alg, REAL x) { return alg == ORC_EXACT ? FN(post_exact)(x) : x; }
static inline REAL FN(sum)(int alg, REAL a, REAL b) {
  if (alg == ORC_EXACT) return a * b;
  if (alg == ORC_LINEAR) return FN(sum_linear)(a, b);
  if (alg == ORC_EXACT_STABLE) return FN(sum_stable)(a, b);
  return FN(sum_min)(a, b);
}

typedef struct {
  const int32_t* row_ptr; const int32_t* col;
  int64_t M, N, K, E;
  int alg, iterations, msg_fmt;
  const double* sf; int sf_rows, sf_cols;
} FN(code_t);

/* One check-node update of every row, in place on c (BpDecoderImpl.cpp:133-159). */
static void FN(check_update)(const FN(code_t)* cd, int64_t t, REAL* c, REAL* tmp) {
  const int alg = cd->alg;
  for (int64_t r = 0; r < cd->M; ++r) {
    REAL* first = c + cd->row_ptr[r];
    int64_t d = cd->row_ptr[r + 1] - cd->row_ptr[r];
    double sf = cd->sf[(t % cd->sf_rows) * cd->sf_cols + (d - 2)];
    double prod = FN(prior)(alg, first[0]);
    for (int64_t j = 1; j < d - 1; ++j) {
      tmp[j] = FN(prior)(alg, first[j]);
      first[j] = (REAL)prod;
      prod = FN(sum)(alg, (REAL)prod, tmp[j]);
    }
    tmp[d - 1] = FN(prior)(alg, first[d - 1]);
    first[d - 1] = (REAL)(sf * FN(post)(alg, (REAL)prod));
    prod = tmp[d - 1];
    for (int64_t j = d - 2; j > 0; --j) {
      first[j] = (REAL)(sf * FN(post)(alg, FN(sum)(alg, first[j], (REAL)prod)));
      prod = FN(sum)(alg, (REAL)prod, tmp[j]);
    }
    first[0] = (REAL)(sf * FN(post)(alg, (REAL)prod));
    if (cd->msg_fmt != ORC_MSG_F32)
      for (int64_t j = 0; j < d; ++j) first[j] = (REAL)orc_quantize((float)first[j], cd->msg_fmt);
  }
}

/* Variable-node update (BpDecoderImpl.cpp:162-179): v2c = prefix + (L + suffix), prefix in
 * ascending edge order from 0, suffix in descending order from L; app = L + all, descending. */
static void FN(bit_update)(const FN(code_t)* cd, const REAL* L, REAL* c, REAL* buf, REAL* app) {
  const int32_t* col = cd->col;
  for (int64_t v = 0; v < cd->N; ++v) app[v] = 0;
  for (int64_t e = 0; e < cd->E; ++e) {
    buf[e] = c[e];
    REAL* ref = &app[col[e]];
    c[e] = *ref;
    *ref += buf[e];
  }
  for (int64_t v = 0; v < cd->N; ++v) app[v] = L[v];
  for (int64_t e = cd->E - 1; e >= 0; --e) {
    REAL* ref = &app[col[e]];
    c[e] += *ref;
    *ref += buf[e];
  }
  if (cd->msg_fmt != ORC_MSG_F32)
    for (int64_t e = 0; e < cd->E; ++e) c[e] = (REAL)orc_quantize((float)c[e], cd->msg_fmt);
}

/* All rows satisfied by hard = (app >= 0)?  Stops at the first failing row. */
static int FN(check_ok)(const FN(code_t)* cd, const REAL* app) {
  for (int64_t r = 0; r < cd->M; ++r) {
    int s = 0;
    for (int32_t e = cd->row_ptr[r]; e < cd->row_ptr[r + 1]; ++e) s ^= (app[cd->col[e]] >= 0.0);
    if (s) return 0;
  }
  return 1;
}

typedef struct {
  const FN(code_t)* cd;
  const double* llr; const double* syst_in; const double* state_in;
  uint8_t* hard; double* msg; double* parity; double* syst; double* state; int32_t* iters;
  uint64_t* lin_side;
  int64_t f0, f1;
} FN(job_t);

static void* FN(run_frames)(void* arg) {
  FN(job_t)* jb = (FN(job_t)*)arg;
  const FN(code_t)* cd = jb->cd;
  const int64_t N = cd->N, K = cd->K, E = cd->E;
  REAL* L = (REAL*)malloc(sizeof(REAL) * N);
  REAL* app = (REAL*)malloc(sizeof(REAL) * N);
  REAL* c = (REAL*)malloc(sizeof(REAL) * E);
  REAL* buf = (REAL*)malloc(sizeof(REAL) * E);
  for (int64_t f = jb->f0; f < jb->f1; ++f) {
    FN(lin_side) = 0xcbf29ce484222325ull;
    const double* llr = jb->llr + f * N;
    for (int64_t v = 0; v < N; ++v) L[v] = (REAL)llr[v];
    if (jb->syst_in)
      for (int64_t v = 0; v < K; ++v) L[v] = (REAL)((double)L[v] + jb->syst_in[f * K + v]);
    if (jb->state_in) {
      for (int64_t e = 0; e < E; ++e) c[e] = (REAL)jb->state_in[f * E + e];
      FN(bit_update)(cd, L, c, buf, app);
    } else {
      for (int64_t e = 0; e < E; ++e) c[e] = L[cd->col[e]];
      if (cd->msg_fmt != ORC_MSG_F32)
        for (int64_t e = 0; e < E; ++e) c[e] = (REAL)orc_quantize((float)c[e], cd->msg_fmt);
    }
    int32_t count = cd->iterations;
    for (int64_t t = 0; t < (int64_t)cd->iterations - 1; ++t) {
      FN(check_update)(cd, t, c, buf);
      FN(bit_update)(cd, L, c, buf, app);
      if (FN(check_ok)(cd, app)) { count = (int32_t)(t + 2); break; }
    }
    FN(check_update)(cd, cd->iterations - 1, c, buf);
    if (jb->iters) jb->iters[f] = count;
    if (jb->lin_side) jb->lin_side[f] = FN(lin_side);
    if (jb->hard) {  /* decode output: app = L, then += c ascending (BpDecoderImpl.cpp:63-69) */
      for (int64_t v = 0; v < N; ++v) app[v] = L[v];
      for (int64_t e = 0; e < E; ++e) app[cd->col[e]] += c[e];
      for (int64_t v = 0; v < K; ++v) jb->hard[f * K + v] = app[v] >= 0;
    }
    if (jb->msg || jb->parity || jb->syst) {  /* soDecode outputs (BpDecoderImpl.cpp:111-129) */
      for (int64_t v = 0; v < N; ++v) app[v] = 0;
      for (int64_t e = 0; e < E; ++e) app[cd->col[e]] += c[e];
      if (jb->syst) for (int64_t v = 0; v < K; ++v) jb->syst[f * K + v] = app[v];
      if (jb->parity) for (int64_t v = 0; v < N; ++v) jb->parity[f * N + v] = app[v];
      if (jb->msg) for (int64_t v = 0; v < K; ++v) jb->msg[f * K + v] = (REAL)(L[v] + app[v]);
    }
    if (jb->state) for (int64_t e = 0; e < E; ++e) jb->state[f * E + e] = c[e];
  }
  free(L); free(app); free(c); free(buf);
  return NULL;
}

static int FN(decode)(const FN(code_t)* cd, const double* llr, int64_t frames,
                      const double* syst_in, const double* state_in,
                      uint8_t* hard, double* msg, double* parity, double* syst, double* state,
                      int32_t* iters, uint64_t* lin_side, int nthreads) {
  FN(lin_init)();
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (frames < nthreads) nthreads = frames > 0 ? (int)frames : 1;
  pthread_t th[256];
  int started[256] = {0};
  FN(job_t) jobs[256];
  int64_t step = (frames + nthreads - 1) / nthreads;
  for (int i = 0; i < nthreads; ++i) {
    int64_t f0 = i * step, f1 = (i + 1) * step < frames ? (i + 1) * step : frames;
    FN(job_t) jb = {cd, llr, syst_in, state_in, hard, msg, parity, syst, state, iters, lin_side, f0, f1 > f0 ? f1 : f0};
    jobs[i] = jb;
  }
  /* chunks 1.. on worker threads, chunk 0 on the caller (src/Codec.h:221-237 fans out alike) */
  for (int i = 1; i < nthreads; ++i)
    started[i] = pthread_create(&th[i], NULL, FN(run_frames), &jobs[i]) == 0;
  FN(run_frames)(&jobs[0]);
  for (int i = 1; i < nthreads; ++i) {
    if (started[i]) pthread_join(th[i], NULL);
    else FN(run_frames)(&jobs[i]);
  }
  return 0;
}

#undef FN
#undef CAT
#undef CAT2
