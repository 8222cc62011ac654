/* CPU oracle for the LDPC belief-propagation decode path: a plain-C restatement of FeCl's
 * decoder (src/detail/BpDecoder/BpDecoderImpl.cpp), its check-node arithmetic
 * (src/detail/LlrMetrics.h, src/detail/LinearTable.h), its syndrome test and scaling table
 * (src/detail/Ldpc.cpp), and a DVB-S2 parity-check construction (src/DvbS2.cpp:1595-1645,
 * corrected for every rate).
 *
 * TEST INFRASTRUCTURE ONLY.  The tests, the smoke check and the benchmark's CPU-baseline leg
 * use it as the checker; the product never links or calls it.
 *
 * Pinning (what makes this restatement trustworthy), per function:
 *   orc_dvbs2_build   r1/2 normal frame byte-identical to FeCl's DvbS2::matrix (golden CSR);
 *                     every rate: E, degree histograms, staircase, H*encode(u) = 0.
 *   orc_dvbs2_encode  uniqueness: the staircase parity solving H*c = 0 is unique, so the
 *                     syndrome test pins it.
 *   orc_bp_decode     prec 64: bit-identical to stock FeCl (golden vectors from oracle/_ref,
 *                     all three algorithms, several iteration counts and scaling factors);
 *                     prec 32: bit-identical to FeCl's own BpDecoderImpl instantiated with a
 *                     float LLR type; iteration counts against counts derived from FeCl.
 *                     Closed forms pin the check node (brute force on tiny rows).
 *   orc_quantize      against numpy float16 and torch bfloat16 round-to-nearest-even.
 */
#ifndef FECL_ORACLE_H
#define FECL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Algorithms (src/DecoderAlgorithm.h:30-34). */
enum { ORC_EXACT = 0, ORC_LINEAR = 1, ORC_APPROXIMATE = 2,
       ORC_EXACT_STABLE = 3 /* BoxSum's value in the LLR domain, for conditioning checks */ };
/* Stored message formats for the FP32 restatement (precision form). */
enum { ORC_MSG_F32 = 0, ORC_MSG_F16 = 1, ORC_MSG_BF16 = 2 };

/* DVB-S2 parity-check matrix: dims = {M, N, K, E}.  Pass NULL arrays to query dims.
 * Returns 0, or -1 for an unknown (n, rate). */
int orc_dvbs2_build(int n, int rate_num, int rate_den, int64_t* dims, int32_t* row_ptr, int32_t* col);

/* Systematic encoder for a staircase code ([info | parity], check r holds p_{r-1}, p_r). */
int orc_dvbs2_encode(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N,
                     const uint8_t* info, uint8_t* cw, int64_t frames);

/* Number of failing rows of H*c (0 = codeword). */
int64_t orc_syndrome_weight(const int32_t* row_ptr, const int32_t* col, int64_t M, const uint8_t* cw);

/* Round-to-nearest-even with saturation to the format's largest finite value. */
float orc_quantize(float x, int fmt);

/* Flooding BP decode of `frames` frames of FeCl-convention LLRs (L = ln p1/p0).
 *   prec      64 (FeCl's double decoder) or 32 (float LLR type)
 *   msg_fmt   ORC_MSG_* storage of both message arrays (only with prec 32)
 *   sf        scaling table [sf_rows][sf_cols], sf_cols = max row degree - 1,
 *             factor for CN update t on a row of degree d = sf[(t % sf_rows)*sf_cols + d-2]
 *   syst_in   optional a-priori on the first K bits  [frames][K]
 *   state_in  optional check-to-variable state        [frames][E]
 * Outputs (each nullable): hard [F][K] (decode), msg [F][K] (soDecode APP), parity [F][N]
 * and syst [F][K] (extrinsic), state [F][E], iters [F] (CN updates executed).
 * Returns 0, or -1 on invalid arguments (iterations < 1, a row of degree < 2). */
int orc_bp_decode(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, int64_t K,
                  const double* llr, int64_t frames, int alg, int iterations,
                  const double* sf, int sf_rows, int sf_cols, int prec, int msg_fmt,
                  const double* syst_in, const double* state_in,
                  uint8_t* hard, double* msg, double* parity, double* syst, double* state,
                  int32_t* iters, int nthreads);

/* orc_bp_decode plus a diagnostic of the Linear rule (alg ORC_LINEAR): lin_side[f]
 * (nullable) receives a 64-bit FNV-1a hash of the side of the table's cutoff (x >= 7, i.e.
 * |a +- b| >= 3.5, where FeCl's correction drops from 0.0298 to 0) that each table lookup of
 * frame f fell on, in lookup order.  A float and a double decode of a frame that take the same
 * number of iterations make the same lookups in the same order, so their hashes differ exactly
 * when some lookup fell on different sides (up to hash collisions). */
int orc_bp_decode_diag(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, int64_t K,
                       const double* llr, int64_t frames, int alg, int iterations,
                       const double* sf, int sf_rows, int sf_cols, int prec, int msg_fmt,
                       const double* syst_in, const double* state_in,
                       uint8_t* hard, double* msg, double* parity, double* syst, double* state,
                       int32_t* iters, uint64_t* lin_side, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
