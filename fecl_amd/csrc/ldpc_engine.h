// Engine internals shared by the C ABI (ldpc_capi.cpp) and the serving loop (ldpc_server.cpp):
// device buffers, cached workspaces, the handle, and the batch runner.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fecl_ldpc.h"
#include "ldpc_code.h"
#include "ldpc_launch.h"

namespace fecl {
namespace engine {

// Records msg as the calling thread's last error and returns status.
int fail(int status, const std::string& msg);

#define HIP_TRY(expr)                                                                            \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return fail(e_ == hipErrorOutOfMemory ? FECL_OUT_OF_MEMORY : FECL_DEVICE_ERROR,            \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                            \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, n);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Pinned host memory (hipHostMalloc), grown on demand.
struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&p, n, hipHostMallocDefault);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// The per-tile buffers of one area of tiles (message rows, channel values, transient hard
// decisions, per-slot state, tile flags and the outputs of frames that finish there).
struct Area {
  DevBuf msg, L, hard, phase, iter, fail, flags, out_hard, out_ext, out_app;
  void release() {
    for (DevBuf* b : {&msg, &L, &hard, &phase, &iter, &fail, &flags, &out_hard, &out_ext, &out_app}) b->release();
  }
};

struct Workspace {
  DevBuf msg, L, hard, phase, iter, fail, flags, active, out_hard, out_ext, out_app;
  // Compaction: the batch's first region of tiles uses the buffers above; later regions
  // alternate between `spare` (three quarters as many tiles) and those buffers.  map[a]: slot -> frame
  // of the region in area a (area 0's first region needs none); src_of: the plan's scratch.
  // Captured graphs compact in place instead: map[] alternates per checkpoint, and ctl holds
  // the device's decision (CompactCtl).
  Area spare;
  DevBuf map[2], src_of, ctl;
  DevBuf stage_in[2], stage_state[2], stage_out[2];   // host path, double-buffered
  HostBuf pin_in[2], pin_state[2], pin_out[2];          // host path: FP32, pinned
  // eager decodes: each step's live frames per tile, written by the control kernel (mapped
  // page-locked memory; host_live_dev is its device address), steps after one another
  int32_t* host_live = nullptr;
  int32_t* host_live_dev = nullptr;
  int64_t host_live_cap = 0;   // entries
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t done = nullptr;   // recorded after the last use, waited on by the next user's stream
  std::vector<std::array<hipEvent_t, 4>> pev;   // per-step timing events (profiling only)
  hipStream_t own_stream = nullptr;
  hipStream_t copy_stream = nullptr;                        // host path: input copies
  hipStream_t out_stream = nullptr;                         // host path: output copies
  hipEvent_t copied[2] = {nullptr, nullptr}, used[2] = {nullptr, nullptr};
  hipEvent_t decoded[2] = {nullptr, nullptr}, drained[2] = {nullptr, nullptr};
  ~Workspace() {
    for (DevBuf* b : {&msg, &L, &hard, &phase, &iter, &fail, &flags, &active, &out_hard, &out_ext, &out_app,
                      &stage_in[0], &stage_in[1], &stage_state[0], &stage_state[1],
                      &stage_out[0], &stage_out[1]})
      b->release();
    spare.release();
    for (DevBuf* b : {&map[0], &map[1], &src_of, &ctl}) b->release();
    for (HostBuf* b : {&pin_in[0], &pin_in[1], &pin_state[0], &pin_state[1],
                       &pin_out[0], &pin_out[1]})
      b->release();
    if (host_live) (void)hipHostFree(host_live);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (done) (void)hipEventDestroy(done);
    for (auto& a : pev)
      for (auto& e : a) (void)hipEventDestroy(e);
    if (own_stream) (void)hipStreamDestroy(own_stream);
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    if (out_stream) (void)hipStreamDestroy(out_stream);
    for (hipEvent_t e : {copied[0], copied[1], used[0], used[1], decoded[0], decoded[1], drained[0], drained[1]})
      if (e) (void)hipEventDestroy(e);
  }
};

}  // namespace engine
}  // namespace fecl

struct fecl_ldpc {
  fecl::Code code;                 // H in generator form, as the decoder uses it
  fecl::GeneratorForm gen;         // encoder blocks of that form (its H is kept in `code`)
  bool has_encoder = true;         // false for a matrix given already in form whose blocks are unknown
  int device = 0;
  fecl::engine::DevBuf d_row_ptr, d_col, d_var_ptr, d_var_edges, d_var_tab, d_sf, d_sf32;
  fecl::DevCode dc{};
  // options
  int32_t algorithm = FECL_APPROXIMATE;
  int32_t iterations = 1;
  int32_t msg_format = FECL_MSG_F32;
  std::vector<double> sf;
  int32_t sf_rows = 1, sf_cols = 1;
  int32_t sf_mode = 0;
  int32_t words = 1;        // device layout: 32-bit words per lane per edge (1, 2 or 4)
  int32_t words_pinned = 0; // set by fecl_ldpc_set_layout / autotune; 0 = automatic
  int32_t vpw = 1;          // variables per wavefront in the variable kernel (1 or 2)
  int32_t vpw_pinned = 0;   // set by fecl_ldpc_set_vars_per_wave / autotune; 0 = automatic
  // Syndrome evaluated inside the next check kernel (one scaling row only) instead of a
  // separate kernel after the variable pass.  Off: the separate kernel reduces per workgroup
  // and sends few atomics, while the fused form puts one device-coherent load per row on the
  // tile's fail word.
  bool fused_syndrome = false;
  int64_t pass_frames = 0;  // frames per pass through the iteration loop; 0 = kMaxPassFrames
  // Compaction of the live frames into fewer tiles once a quarter of a region's tiles could go
  // (eager decodes without the fused syndrome; see run_batch).
  bool compaction = true;
  // Captured multi-code graphs run as one persistent kernel (ldpc_persist.hip) when every
  // handle of the graph asks for it and is eligible; otherwise (default) as recorded per-step
  // launches, which are faster on the seven-rate batch (the persistent kernel holds every
  // phase's registers at once, 2 waves per SIMD).
  bool persistent = false;
  int32_t cache_policy = -1; // message-row cache hints: -1 by working set, else DevOpts::nt_rows bits
  std::mutex mu;
  std::vector<std::unique_ptr<fecl::engine::Workspace>> pool;
  // kernel timing (fecl_ldpc_profile): check, variable, syndrome+control
  bool profiling = false;
  double prof_ms[3] = {0, 0, 0};
  int64_t prof_launches[3] = {0, 0, 0};
  ~fecl_ldpc() {
    pool.clear();
    for (fecl::engine::DevBuf* b : {&d_row_ptr, &d_col, &d_var_ptr, &d_var_edges, &d_var_tab, &d_sf, &d_sf32})
      b->release();
  }
};

namespace fecl {
namespace engine {

struct BatchIO {
  const void* llr;        // [F][N]
  const void* syst;       // [F][K] or null
  const void* state;      // [F][E] or null
  bool dbl_in;
  BatchResults res;
  bool dbl_out;
};

// The handle's kernel options with the scaling tables in device memory of its own, for users
// that keep running on the options a handle had when they were made (captured graphs, the
// server): fecl_ldpc_set_options rewrites or reallocates the handle's tables.
struct OptsCopy {
  DevBuf sf, sf32;
  DevOpts opts{};
  ~OptsCopy() {
    sf.release();
    sf32.release();
  }
};
int copy_opts(const fecl_ldpc* h, OptsCopy* out);

int check_handle(const fecl_ldpc* h);
// DevOpts::nt_rows for a batch of `tiles` tiles of h's layout: the handle's policy, or with
// the automatic policy non-temporal rows in both kernels once the batch's rows and channel
// values exceed kNtBatchBytes (2.5x the 256 MB Infinity Cache, so little of them stays cached
// from one kernel to the next anyway).  Measured on DVB-S2 r1/2 and r3/4: from ~0.7 GB up
// the hints gain 2-5 %; at 0.3-0.6 GB they lose up to 7 % in the check kernel.
constexpr int64_t kNtBatchBytes = int64_t(640) << 20;
int32_t nt_rows_for(const fecl_ldpc* h, const Layout& lay, int64_t tiles);
Layout layout_of(const fecl_ldpc* h);
DevOpts make_opts(const fecl_ldpc* h);
// Sizes ws for `frames` frames of h's layout and returns the device views.
int prepare(fecl_ldpc* h, Workspace* ws, const Layout& lay, int64_t frames, bool soft, DevState* sp, DevOut* op);
// Frames per pass unless fecl_ldpc_set_pass_frames says otherwise.  A pass runs every iteration
// of its frames; the variable update starts on the tiles the check update wrote last, which the
// 256 MB Infinity Cache still holds, and with four 256-frame tiles a quarter of the rows come
// from there (4096-frame batches: 0.929 -> 0.959 Gb/s FP32, 1.742 -> 1.766 FP16 against one
// pass; 2048-frame passes 0.932 and 1.730).
constexpr int64_t kMaxPassFrames = 1024;
// A compaction is enqueued only while at least this many steps may remain.
constexpr int kCompactMinSteps = 4;
// Captured graphs: a device-decided compaction after every this many steps (3, 4 and 6 measured
// on the uneven seven-rate batch: 3.37, 3.40 and 3.37 Gb/s).
constexpr int kGraphCompactEvery = 4;
// Sizes ws's in-place compaction buffers for a captured decode of `frames` frames (before
// capture: nothing may allocate while a graph is recorded).
int prepare_graph_compaction(fecl_ldpc* h, Workspace* ws, int64_t frames);
// Decodes frames [frame0, frame0 + frames) of io in one pass (see ldpc_capi.cpp).  base
// (nullable): kernel options to use instead of make_opts(h) (their scaling tables).
int run_batch(fecl_ldpc* h, Workspace* ws, hipStream_t st, const BatchIO& io, int64_t frames, int64_t frame0,
              bool capture, const DevOpts* base = nullptr);
// Decodes frames [frame0, frame0 + frames) of io in consecutive passes of h->pass_frames.
int run_passes(fecl_ldpc* h, Workspace* ws, hipStream_t st, const BatchIO& io, int64_t frames, int64_t frame0,
               bool capture, const DevOpts* base = nullptr);

// Persistent multi-code decode (ldpc_persist.hip).  eligible: whether these handles can share
// one persistent launch, and their common layout.  create: sizes the batch's workspaces and
// control block (before capture).  enqueue: one decode (counter reset, ingests, the kernel),
// capturable.  error: whether the last finished launch ended on its watchdog.
struct PersistBatch;
bool persist_eligible(fecl_ldpc* const* handles, int32_t count, Layout* common);
int persist_create(fecl_ldpc* const* handles, const int64_t* frames, int32_t count, const float* const* llr,
                   uint8_t* const* msg, int32_t* const* iterations, const Layout& lay, const DevOpts* opts,
                   PersistBatch** out);
int persist_enqueue(PersistBatch* b, fecl_ldpc* const* handles, hipStream_t st);
int persist_error(PersistBatch* b, int32_t* error);
void persist_destroy(PersistBatch* b);

}  // namespace engine
}  // namespace fecl
