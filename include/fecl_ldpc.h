/* fecl_ldpc.h — C ABI of the MI355X LDPC belief-propagation decoder.
 *
 * Drop-in boundary for FeCl's LDPC decode path: the functions below are what
 * fec::Ldpc::decodeBlocks / soDecodeBlocks (src/Ldpc.h:102-103, src/Ldpc.cpp:49-59) and the
 * public surface around them (src/Ldpc.h:59-108, src/Codec.h:60-82, src/detail/Ldpc.h:56-74)
 * bind to.  Plain pointers and sizes only; no exceptions cross this boundary: every call
 * returns a fecl_status and fecl_last_error() holds the message (thread-local).
 *
 * Conventions (FeCl's):
 *   - LLR L = ln p(x=1|y) / p(x=0|y) (src/Codec.h:204-205); hard decision bit = (L >= 0).
 *   - Frames are contiguous: LLR [frames][N], hard output [frames][K] (one byte per bit,
 *     0 or 1, in place of BitField<size_t>), state [frames][E] in the parity-check matrix's
 *     row-major edge order.
 *   - A code is H in CSR form (rows = checks); K = N - M (src/detail/Ldpc.cpp:46-55).
 */
#ifndef FECL_LDPC_H
#define FECL_LDPC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  FECL_OK = 0,
  FECL_INVALID_ARGUMENT = 1, /* FeCl throws std::invalid_argument here */
  FECL_OUT_OF_MEMORY = 2,
  FECL_DEVICE_ERROR = 3,
  FECL_NOT_SUPPORTED = 4
} fecl_status;

/* fec::DecoderAlgorithm (src/DecoderAlgorithm.h:30-34). */
typedef enum { FECL_EXACT = 0, FECL_LINEAR = 1, FECL_APPROXIMATE = 2 } fecl_algorithm;

/* Storage format of the edge messages on the device.  Check-node and variable-node
 * arithmetic is FP32 in every format; 16-bit formats saturate at the format's largest
 * finite value and round to nearest even. */
typedef enum { FECL_MSG_F32 = 0, FECL_MSG_F16 = 1, FECL_MSG_BF16 = 2 } fecl_msg_format;

/* Ldpc::DecoderOptions (src/detail/Ldpc.h:56-74), flattened.  `scaling` is the table FeCl
 * builds from its {degree -> factors} map (Structure::scalingMapToVector,
 * src/detail/Ldpc.cpp:64-112): [scaling_rows][scaling_cols], scaling_cols = max row
 * degree - 1, factor for check update t on a row of degree d = scaling[(t % rows)*cols + d-2].
 * NULL means 1.0 everywhere.  Build one from a map with fecl_ldpc_scaling_table(). */
typedef struct {
  int32_t algorithm;      /* fecl_algorithm; FeCl's default is FECL_APPROXIMATE */
  int32_t iterations;     /* >= 1 */
  const double* scaling;  /* may be NULL */
  int32_t scaling_rows;   /* 1 or iterations */
  int32_t scaling_cols;   /* max row degree - 1 */
  int32_t msg_format;     /* fecl_msg_format */
} fecl_decoder_options;

typedef struct fecl_ldpc fecl_ldpc;

/* Message of the last failed call on this thread ("" if none). */
const char* fecl_last_error(void);

/* Ldpc::DvbS2::matrix(n, rate) (src/Ldpc.h:66-75, src/DvbS2.cpp:1595-1645), correct for every
 * normal (64800) and short (16200) DVB-S2 rate.  dims = {M, N, K, E}; pass NULL row_ptr/col to
 * query sizes (row_ptr has M+1 entries, col has E). */
int fecl_dvbs2_matrix(int32_t n, int32_t rate_num, int32_t rate_den, int64_t* dims,
                      int32_t* row_ptr, int32_t* col);

/* Structure::scalingMapToVector (src/detail/Ldpc.cpp:64-112): map entry i is
 * degree keys[i] (0 = default) -> values[offsets[i] .. offsets[i]+lengths[i]).
 * Writes rows*cols doubles into `table` (capacity `table_capacity`). */
int fecl_ldpc_scaling_table(const int32_t* row_ptr, int64_t M, int64_t iterations,
                            const int64_t* keys, const int64_t* lengths, const int64_t* offsets,
                            const double* values, int64_t entries,
                            double* table, int64_t table_capacity, int32_t* rows, int32_t* cols);

/* Ldpc::Gallager::matrix(n, wc, wr, seed) (src/Ldpc.h:63-65, src/Ldpc.cpp:77-100): the same
 * matrix as FeCl for the same arguments.  dims = {M, N, E}; pass NULL row_ptr/col to query the
 * sizes (row_ptr has M+1 entries, col has E). */
int fecl_gallager_matrix(int64_t n, int64_t wc, int64_t wr, uint64_t seed, int64_t* dims,
                         int32_t* row_ptr, int32_t* col);

/* The matrix FeCl decodes and encodes with for a given H: Structure(EncoderOptions(H)).checks()
 * and msgSize() (src/detail/Ldpc.cpp:46-55, 216-331).  Rows and columns of H are permuted into
 * [A | B | T] over [C | D | E] with T lower triangular; a codeword is then [message | parity],
 * and message bit k is codeword bit k of this matrix.  K exceeds N - M by the number of
 * dependent checks.  H is unchanged for staircase codes such as DVB-S2.  dims = {M, N, K, E};
 * pass NULL out arrays to query the sizes.  Columns must be ascending within each row. */
int fecl_ldpc_structure(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N,
                        int64_t* dims, int32_t* out_row_ptr, int32_t* out_col);

/* Codec::encode (src/Codec.h:68-72, src/detail/Ldpc.cpp:180-209) for the code of H without a
 * decoder handle or a device: msg [frames][K] in, codewords [frames][N] out (0/1 bytes), K and
 * the bit order as fecl_ldpc_structure gives them. */
int fecl_encode(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, const uint8_t* msg,
                int64_t frames, uint8_t* parity);

/* fec::Ldpc(EncoderOptions(H), DecoderOptions) (src/Ldpc.h:82-84): puts H in FeCl's form
 * (fecl_ldpc_structure) and copies that to the device.  Columns must be ascending within each
 * row. */
int fecl_ldpc_create(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N,
                     const fecl_decoder_options* options, fecl_ldpc** out);
int fecl_ldpc_destroy(fecl_ldpc* h);

/* Ldpc::setDecoderOptions / getDecoderOptions (src/Ldpc.h:91-92).  get copies the scaling
 * table into `scaling` (capacity `capacity` doubles) when non-NULL. */
int fecl_ldpc_set_options(fecl_ldpc* h, const fecl_decoder_options* options);
int fecl_ldpc_get_options(const fecl_ldpc* h, fecl_decoder_options* options, double* scaling,
                          int64_t capacity);

/* fec::Ldpc(const Structure&) (src/Ldpc.h:83) for a binding inside FeCl: a decoder for a matrix
 * already in FeCl's form (structure().checks(), msgSize() = K), used as given, without the
 * transform.  Such a handle encodes and saves only when the matrix is its own form (staircase
 * codes such as DVB-S2); otherwise FeCl keeps encoding with its Structure. */
int fecl_ldpc_create_checks(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, int64_t K,
                            const fecl_decoder_options* options, fecl_ldpc** out);

/* Saved codecs: fec::detail::save / load of a fec::Ldpc (src/Serialization.h:81-115), in FeCl's
 * format: a Boost 1.58 binary archive of a `const fec::Codec*` with the types Convolutional,
 * Turbo and Ldpc registered in that order, as FeCl's Matlab wrapper writes them
 * (wrap/wrap.h:35, wrap/Matlab/util/SaveLoad.h:40-64).  The archive holds the structure (the
 * matrix in generator form and its encoder blocks), the decoder options and the work group
 * size; the message format is not part of it (loaded handles use FP32 messages).
 *
 * fecl_ldpc_save writes the handle's archive (*size = its length; buffer NULL queries the
 * length).  fecl_ldpc_load makes a handle from an archive, without repeating the transform;
 * bytes after the archive are ignored.  *work_group_size (nullable) receives the stored value. */
int fecl_ldpc_save(const fecl_ldpc* h, int32_t work_group_size, uint8_t* buffer, int64_t capacity,
                   int64_t* size);
int fecl_ldpc_load(const uint8_t* buffer, int64_t size, fecl_ldpc** out, int32_t* work_group_size);
/* The same archives without a decoder handle or a device: fecl_archive_save writes the archive
 * of fec::Ldpc(EncoderOptions(H), options, work_group_size); fecl_archive_load reads one back
 * into the stored matrix (dims = {M, N, K, E}; NULL arrays query the sizes), options (the
 * scaling table is copied into `scaling`, capacity doubles, when non-NULL) and work group size. */
int fecl_archive_save(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N,
                      const fecl_decoder_options* options, int32_t work_group_size, uint8_t* buffer,
                      int64_t capacity, int64_t* size);
int fecl_archive_load(const uint8_t* buffer, int64_t size, int64_t* dims, int32_t* row_ptr,
                      int32_t* col, fecl_decoder_options* options, double* scaling,
                      int64_t capacity, int32_t* work_group_size);

/* Codec::msgSize/systSize/paritySize/stateSize (src/Codec.h:60-63). */
int fecl_ldpc_sizes(const fecl_ldpc* h, int64_t* msg, int64_t* syst, int64_t* parity,
                    int64_t* state);
/* Structure::checks() (src/detail/Ldpc.h:115): the handle's matrix in FeCl's form, whose edge
 * order is the state order.  dims = {M, N, K, E}; NULL arrays query the sizes. */
int fecl_ldpc_checks(const fecl_ldpc* h, int64_t* dims, int32_t* row_ptr, int32_t* col);

/* Codec::encode / Codec::check (src/Codec.h:68-72; src/detail/Ldpc.cpp:160-209), host
 * computation, any code: msg [frames][K] and parity [frames][N] are 0/1 bytes. */
int fecl_ldpc_encode(const fecl_ldpc* h, const uint8_t* msg, int64_t frames, uint8_t* parity);
int fecl_ldpc_check(const fecl_ldpc* h, const uint8_t* parity, int64_t frames, int32_t* ok);
/* Codec::encode on the device (same results as fecl_ldpc_encode): msg [frames][K] and parity
 * [frames][N] are device buffers of 0/1 bytes; enqueued on `stream` (a hipStream_t, NULL =
 * the null stream).  Staircase codes with N <= 65536. */
int fecl_ldpc_encode_device(fecl_ldpc* h, const uint8_t* msg, int64_t frames, uint8_t* parity, void* stream);

/* Codec::decode -> Ldpc::decodeBlocks (src/Codec.h:210-238, src/Ldpc.cpp:55-59), host buffers:
 * llr [frames][N] double in, msg [frames][K] (0/1 bytes) out.  iterations (nullable)
 * receives the per-frame count of check-node updates executed (FeCl computes it but does not
 * expose it: src/detail/BpDecoder/BpDecoderImpl.cpp:48-61). */
int fecl_ldpc_decode(fecl_ldpc* h, const double* llr, int64_t frames, uint8_t* msg,
                     int32_t* iterations);

/* Codec::soDecode -> Ldpc::soDecodeBlocks (src/Codec.h:249-307), host buffers.
 * Inputs: parity [F][N] (required), syst [F][K] and state [F][E] (nullable).
 * Outputs (each nullable): msg [F][K] a-posteriori, syst [F][K] and parity [F][N]
 * extrinsic, state [F][E] check-to-variable messages. */
int fecl_ldpc_sodecode(fecl_ldpc* h, const double* parity_in, const double* syst_in,
                       const double* state_in, int64_t frames, double* msg_out,
                       double* syst_out, double* parity_out, double* state_out,
                       int32_t* iterations);

/* Device-resident forms for callers that keep data in HBM (same semantics, FP32 LLRs,
 * asynchronous on `stream` (a hipStream_t; NULL = the null stream)).  All pointers are device
 * pointers; nullable ones as above. */
int fecl_ldpc_decode_device(fecl_ldpc* h, const float* llr, int64_t frames, uint8_t* msg,
                            int32_t* iterations, void* stream);
int fecl_ldpc_sodecode_device(fecl_ldpc* h, const float* parity_in, const float* syst_in,
                              const float* state_in, int64_t frames, float* msg_out,
                              float* syst_out, float* parity_out, float* state_out,
                              int32_t* iterations, void* stream);

/* fecl_ldpc_decode_device without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device or host memory during the call (code recorded into a
 * hipGraph, serving loops that forbid allocation).  The batch's workspace is a fixed 128 MiB
 * array compiled into the library (one per device, held from the first call until the process
 * ends); the call makes no hipMalloc/hipFree/hipHostMalloc and creates no stream or event.
 * Same results as fecl_ldpc_decode_device.  It runs every one of the handle's `iterations`
 * steps (no host reads the live count; steps after the last frame finishes return at once),
 * with one 32-bit word per lane per edge (64-frame tiles for FP32 messages, 128 for 16-bit),
 * no compaction and no passes: a batch whose workspace exceeds the array is refused with
 * FECL_INVALID_ARGUMENT (DVB-S2 n = 64800: 64 frames with FP32 messages, 128 with 16-bit).
 * Calls that overlap on one device (other streams or threads) race on the array. */
int fecl_ldpc_decode_device_noalloc(fecl_ldpc* h, const float* llr, int64_t frames, uint8_t* msg,
                                    int32_t* iterations, void* stream);

/* Kernel timing (no counterpart in the reference; instrumentation for benchmarks).
 * While enabled, each decode records HIP events on its launch stream around the check-node,
 * variable-node and syndrome/control launches of every step and waits for the batch to finish.
 * profile_read returns the accumulated milliseconds and launch counts of those three groups.
 * Enabling resets the totals. */
int fecl_ldpc_profile(fecl_ldpc* h, int32_t enable);
int fecl_ldpc_profile_read(fecl_ldpc* h, double* ms /*[3]*/, int64_t* launches /*[3]*/);

/* Device layout (no counterpart in the reference).  Frames are decoded in tiles of
 * 64 * words * (1 for FP32, 2 for 16-bit) frames (the frames one workgroup serves); each GPU
 * lane moves `words` 32-bit words per edge.  set_layout pins words (1, 2, 4; 0 = widest that
 * fits the code's degrees).  set_vars_per_wave pins how many variables one wavefront of the
 * variable kernel serves (1, 2; 0 = automatic, 1); 2 is refused (FECL_NOT_SUPPORTED) where
 * it exceeds the register budget of the current layout, and falls back to 1 if a later
 * layout change makes it not fit.  autotune times `probe_iterations` iterations of a
 * `frames`-frame decode for every (words, variables per wave) pair that fits, pins the
 * fastest and returns its words; ms[6] (nullable) receives the times in the order
 * (1,1) (1,2) (2,1) (2,2) (4,1) (4,2) (-1 = not eligible).  None of these may run
 * concurrently with a decode on the same handle. */
int fecl_ldpc_set_layout(fecl_ldpc* h, int32_t words);
int fecl_ldpc_set_vars_per_wave(fecl_ldpc* h, int32_t vars);
int fecl_ldpc_get_vars_per_wave(const fecl_ldpc* h, int32_t* vars);
/* Syndrome schedule (no counterpart in the reference; results are identical either way):
 * 0 (default) = a separate kernel after each variable pass; 1 = folded into the next check
 * kernel (used only when the scaling factors do not vary with the iteration). */
int fecl_ldpc_set_fused_syndrome(fecl_ldpc* h, int32_t enable);
int fecl_ldpc_get_layout(const fecl_ldpc* h, int32_t* words, int32_t* frames_per_tile);
/* Frames per pass (no counterpart in the reference; results are identical for any value).
 * A batch is decoded in consecutive passes of at most this many frames (rounded up to whole
 * tiles), each pass running all of its iterations before the next starts, so that a pass's
 * messages and channel values can stay in the device's last-level cache between
 * iterations.  0 (default) = 1024 frames per pass. */
int fecl_ldpc_set_pass_frames(fecl_ldpc* h, int64_t frames);
/* Compaction (no counterpart in the reference; results are identical either way).  A tile
 * of frames works until its last frame finishes, so once frames converge most tiles carry a
 * few live frames.  With compaction on (default), an eager decode moves the live frames into
 * three quarters as many tiles or fewer whenever they fit and enough iterations remain, and
 * later iterations run on those tiles only; the batch's workspace grows by up to 75 %.
 * Captured decodes (fecl_ldpc_graph_create) compact in place instead: after every fourth
 * step, while at least four remain, a device kernel decides whether the live frames fit in
 * fewer tiles and moves them, in one pass, into the region's first tiles; such a graph holds
 * three more slot-sized maps.  The fused syndrome never compacts. */
int fecl_ldpc_set_compaction(fecl_ldpc* h, int32_t enable);
/* Cache hints for the message rows (no counterpart in the reference; results are identical
 * for any value): bit 0 = non-temporal loads/stores in the check kernel, bit 1 = in the
 * variable kernel; -1 (default) = both once a batch's rows and channel values exceed 640 MiB,
 * neither below (smaller batches gain from rows that stay in the Infinity Cache). */
int fecl_ldpc_set_cache_policy(fecl_ldpc* h, int32_t policy);
int fecl_ldpc_autotune(fecl_ldpc* h, int64_t frames, int32_t probe_iterations, int32_t* best_words,
                       double* ms /*[6]*/);
/* Autotune cache (no counterpart in the reference).  fecl_ldpc_autotune keeps each pick under
 * (GPU architecture and compute-unit count, code: sizes and a hash of H, algorithm, message
 * format, frames rounded up to a multiple of 64 up to 8192 frames, to a power of two above);
 * a later autotune with the same key pins the kept pick without probing (ms[] then holds the
 * kept probe times).  mode 0 = off, 1 = this
 * process (the default), 2 = this process and the text file `path`, read now and appended with
 * every new pick.  Without a config call, the environment variable FECL_AUTOTUNE_CACHE names
 * such a file at the first autotune.  stats: lookups that hit and missed, picks held, file
 * lines skipped as malformed, and appends that failed (a pick that cannot be written to the
 * file is still pinned and kept in the process).  clear forgets the picks held in this process
 * (not the file).  Any argument of stats may be NULL. */
int fecl_autotune_cache_config(int32_t mode, const char* path);
int fecl_autotune_cache_stats(int64_t* hits, int64_t* misses, int64_t* entries, int64_t* skipped_lines,
                              int64_t* write_failures);
int fecl_autotune_cache_clear(void);

/* Multi-code batch in one launch (no counterpart in the reference: FeCl decodes one code per
 * call).  Captures, for each i < count, a full decode of frames[i] frames of code handles[i]
 * (device LLRs llr[i] [frames[i]][N_i] FP32 -> device msg[i] [frames[i]][K_i] bytes and
 * iterations[i] [frames[i]], nullable) into one hipGraph.  Each launch decodes whatever the
 * input buffers hold at that time.  The options (and layout) of each handle are those at
 * creation time.
 *
 * Two forms, with the same results.  Persistent (chosen when every handle asked for it with
 * fecl_ldpc_set_graph_persistent and all are min-sum without scaling factors in one message
 * format and lane layout): a counter reset, each code's ingest, and ONE persistent
 * kernel whose workgroups take check, variable and syndrome items of all codes from one work
 * queue, with per-tile counters in place of kernel boundaries; converged tiles drop out, and
 * live frames are compacted into the other of two areas at device-decided checkpoints
 * (workspace: two areas of each code's tiles).  Otherwise (the default, faster on the DVB-S2
 * seven-rate batch): every step of every code is recorded as its own launches on concurrent
 * branches; frames that finish early leave the remaining steps as cheap no-ops.  fecl_ldpc_graph_check (after a launch has finished)
 * reports FECL_DEVICE_ERROR if a persistent launch stopped on its watchdog (a wait that did not
 * end within seconds); fecl_ldpc_graph_is_persistent tells which form a graph has. */
typedef struct fecl_ldpc_graph fecl_ldpc_graph;
int fecl_ldpc_graph_create(fecl_ldpc* const* handles, const int64_t* frames, int32_t count,
                           const float* const* llr, uint8_t* const* msg, int32_t* const* iterations,
                           fecl_ldpc_graph** out);
int fecl_ldpc_graph_launch(fecl_ldpc_graph* g, void* stream);
int fecl_ldpc_graph_check(fecl_ldpc_graph* g);
int fecl_ldpc_graph_is_persistent(const fecl_ldpc_graph* g, int32_t* persistent);
int fecl_ldpc_graph_destroy(fecl_ldpc_graph* g);
/* 1: graphs made with this handle may use the persistent form (default 0: per-step form). */
int fecl_ldpc_set_graph_persistent(fecl_ldpc* h, int32_t enable);

/* Continuous-batching decode server (no counterpart in the reference).  Each of the `count`
 * (at most 32) codes gets a pool of device frame slots: slots[i] frames, or, where slots is NULL or
 * slots[i] <= 0, a share of memory_bytes (<= 0: half of the device's free memory) sized by one
 * slot's footprint.  A worker thread repeats rounds: admit queued frames into free slots, run
 * check/variable/syndrome steps (fecl_ldpc_server_set_round_steps, default 1) on every
 * pool holding live frames, retire finished frames.  Requests arriving while others run join
 * at the next round; each frame's result equals a batch decode of it (frames share no state).
 * Handles must outlive the server; each pool decodes with the options (and layout) its handle
 * had when the server was created.
 *   submit: LLRs [frames][N] (FP32) of code `code`, in host memory (pageable or page-locked)
 *           or in device memory of the server's device (read in place; the caller's writes
 *           must be complete when submit is called); msg [frames][K] (and iterations
 *           [frames], nullable) receive the results in host memory; all three stay valid
 *           until wait() reports done.
 *   wait:   blocks up to timeout_s (< 0: forever); done = 1 and latency_s = completion -
 *           submission time once finished (the ticket is then released). */
typedef struct fecl_ldpc_server fecl_ldpc_server;
int fecl_ldpc_server_create(fecl_ldpc* const* handles, int32_t count, const int64_t* slots,
                            int64_t memory_bytes, fecl_ldpc_server** out);
int fecl_ldpc_server_slots(const fecl_ldpc_server* s, int32_t code, int64_t* slots);
int fecl_ldpc_server_submit(fecl_ldpc_server* s, int32_t code, const float* llr, int64_t frames,
                            uint8_t* msg, int32_t* iterations, uint64_t* ticket);
int fecl_ldpc_server_wait(fecl_ldpc_server* s, uint64_t ticket, double timeout_s, int32_t* done,
                          double* latency_s);
int fecl_ldpc_server_stats(fecl_ldpc_server* s, int64_t* frames_done, int64_t* steps,
                           int64_t* queued_frames);
/* Steps per round (1..64; default 1), from the next round on.  More steps cost fewer host
 * round trips per step, but a finished frame waits up to steps - 1 more steps for its
 * retirement and admissions come less often. */
int fecl_ldpc_server_set_round_steps(fecl_ldpc_server* s, int32_t steps);
int fecl_ldpc_server_destroy(fecl_ldpc_server* s);

/* Ldpc::puncturing (src/Ldpc.h:94, src/detail/Ldpc.cpp:333-353): the index sequence of FeCl's
 * Permutation for PunctureOptions().mask(mask).systMask(syst_mask) (0/1 bytes; a length of 0
 * is an empty mask), for a code of syst_size systematic bits in parity_size.  The
 * permutation's input size is parity_size; sequence (capacity entries, nullable) receives
 * its `count` indices, exactly as FeCl forms them.  Host computation. */
int fecl_ldpc_puncturing(int64_t syst_size, int64_t parity_size, const uint8_t* mask, int64_t mask_len,
                         const uint8_t* syst_mask, int64_t syst_mask_len, int64_t* sequence, int64_t capacity,
                         int64_t* count);
/* Permutation::permute / dePermute (src/Permutation.h) of `frames` FP32 blocks on the device:
 * permute (inverse = 0) out[f][i] = in[f][sequence[i]], blocks of input_size -> output_size;
 * dePermute (inverse = 1) out[f][sequence[i]] = in[f][i], other entries 0, blocks of
 * output_size -> input_size (so punctured positions enter the decoder as 0-LLR).
 * sequence [output_size] is a device buffer of indices < input_size without repeats (as
 * puncturing produces); enqueued on `stream` (a hipStream_t, NULL = the null stream). */
int fecl_permute_device(const int64_t* sequence, int64_t output_size, int64_t input_size, const float* in,
                        float* out, int64_t frames, int32_t inverse, void* stream);

/* Releases the cached device workspaces of a handle (they are reused across calls). */
int fecl_ldpc_trim(fecl_ldpc* h);

#ifdef __cplusplus
}
#endif

#endif /* FECL_LDPC_H */
