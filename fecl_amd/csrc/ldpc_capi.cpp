// C ABI of the decoder (include/fecl_ldpc.h): handles, batch engine, host/device entry points.
//
// Reference behaviour mirrored here:
//   size validation and errors     src/Codec.h:140-143, 210-216, 249-270
//   options / scaling table        src/detail/Ldpc.h:56-74, src/detail/Ldpc.cpp:57-112, 134-138
//   encode / check                 src/detail/Ldpc.cpp:160-209
//   decode / soDecode semantics    src/detail/BpDecoder/BpDecoderImpl.cpp:38-130
// Frames are independent, so a call splits its frames into chunks that fit a cached device
// workspace and decodes each chunk as one batch of tiles (no cross-frame arithmetic).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <array>
#include <cmath>
#include <cstring>
#include <future>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/fecl_ldpc.h"
#include "ldpc_archive.h"
#include "ldpc_code.h"
#include "ldpc_engine.h"
#include "ldpc_host_convert.h"
#include "ldpc_launch.h"
#include "ldpc_tune_cache.h"

namespace fecl {
namespace engine {

thread_local std::string g_err;

int fail(int status, const std::string& msg) {
  g_err = msg;
  return status;
}



// Lane layout (unless one was pinned): the first that the kernels allow for this code and
// options, in the order measured fastest on the DVB-S2 workloads (FP32: 4, 2, 1 words per
// lane; 16-bit: 2, 1, 4), preferring layouts whose check kernel holds a row in registers.
// Then the variables per wavefront (pinned, else 1) where they fit that layout.
void choose_layout(fecl_ldpc* h) {
  const int order32[3] = {4, 2, 1}, order16[3] = {2, 1, 4};
  const int* cand = h->msg_format == FECL_MSG_F32 ? order32 : order16;
  const Code& c = h->code;
  h->words = 0;
  if (h->words_pinned) {
    Layout lay;
    lay.fmt = h->msg_format;
    lay.words = h->words_pinned;
    if (layout_fits(lay, h->algorithm, c.max_row_degree, c.max_var_degree)) h->words = h->words_pinned;
  }
  for (int pass = 0; pass < 2 && !h->words; ++pass) {
    for (int i = 0; i < 3 && !h->words; ++i) {
      Layout lay;
      lay.fmt = h->msg_format;
      lay.words = cand[i];
      if (!layout_fits(lay, h->algorithm, c.max_row_degree, c.max_var_degree)) continue;
      if (pass == 0 && !cn_register_staged(lay, h->algorithm, h->sf_mode, c.max_row_degree)) continue;
      h->words = cand[i];
    }
  }
  if (!h->words) h->words = 1;
  Layout lay;
  lay.fmt = h->msg_format;
  lay.words = h->words;
  lay.vpw = h->vpw_pinned ? h->vpw_pinned : 1;
  h->vpw = vpw_fits(lay, h->code.max_var_degree) ? lay.vpw : 1;
}

int set_options(fecl_ldpc* h, const fecl_decoder_options* o) {
  if (!o) return fail(FECL_INVALID_ARGUMENT, "null options");
  if (o->algorithm < 0 || o->algorithm > 2) return fail(FECL_INVALID_ARGUMENT, "Invalid decoder algorithm");
  if (o->iterations < 1) return fail(FECL_INVALID_ARGUMENT, "iterations must be >= 1");
  if (o->msg_format < 0 || o->msg_format > 2) return fail(FECL_INVALID_ARGUMENT, "Invalid message format");
  const int32_t cols = h->code.max_row_degree - 1;
  std::vector<double> sf;
  int32_t rows = 1;
  if (o->scaling) {
    if (o->scaling_cols != cols) return fail(FECL_INVALID_ARGUMENT, "Wrong size for scaling factor");
    if (o->scaling_rows != 1 && o->scaling_rows != o->iterations)
      return fail(FECL_INVALID_ARGUMENT, "Wrong size for scaling factor");
    rows = o->scaling_rows;
    sf.assign(o->scaling, o->scaling + (size_t)rows * cols);
  } else {
    sf.assign((size_t)cols, 1.0);
  }
  if (o->algorithm == FECL_APPROXIMATE && h->code.max_row_degree > 64)
    return fail(FECL_NOT_SUPPORTED, "min-sum check rows are limited to 64 edges");
  if (o->algorithm != FECL_APPROXIMATE && h->code.max_row_degree > 32)
    return fail(FECL_NOT_SUPPORTED, "Exact/Linear check rows are limited to 32 edges");
  int mode = 0;
  std::vector<float> sf32(sf.size());
  for (size_t i = 0; i < sf.size(); ++i) {
    sf32[i] = (float)sf[i];
    if (sf[i] != 1.0) mode = std::max(mode, 1);
    if ((double)sf32[i] != sf[i]) mode = 2;
  }
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(h->d_sf.ensure(sf.size() * sizeof(double)));
  HIP_TRY(h->d_sf32.ensure(sf.size() * sizeof(float)));
  HIP_TRY(hipMemcpy(h->d_sf.p, sf.data(), sf.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_sf32.p, sf32.data(), sf32.size() * sizeof(float), hipMemcpyHostToDevice));
  h->algorithm = o->algorithm;
  h->iterations = o->iterations;
  h->msg_format = o->msg_format;
  h->sf = std::move(sf);
  h->sf_rows = rows;
  h->sf_cols = cols;
  h->sf_mode = mode;
  choose_layout(h);
  return FECL_OK;
}

DevOpts make_opts(const fecl_ldpc* h) {
  DevOpts o{};
  o.sf = h->d_sf.as<double>();
  o.sf32 = h->d_sf32.as<float>();
  o.sf_rows = h->sf_rows;
  o.sf_cols = h->sf_cols;
  o.sf_mode = h->sf_mode;
  o.iterations = h->iterations;
  o.algorithm = h->algorithm;
  // With one scaling row the check update after a passing syndrome equals FeCl's final one
  // (factor row I-1), so the syndrome can ride in the next check kernel.
  o.fused_syn = (h->fused_syndrome && h->sf_rows == 1) ? 1 : 0;
  // FeCl's Linearlog1pexpm table in float: y[i] = log(1 + exp(-i/2)) evaluated in double
  // (src/detail/LinearTable.h:40-44, 63), slopes y[i+1] - y[i] in float.
  float y[8];
  for (int i = 0; i < 8; ++i) y[i] = (float)std::log(1.0 + std::exp(-((double)(float)i / (double)2.0f)));
  for (int i = 0; i < 8; ++i) {
    // knot 7 enters only through the slope of segment 6: from x = 7 on FeCl's f is 0, which
    // the kernels read as the (0, 0) entry 7
    o.lin_y[i] = i < 7 ? y[i] : 0.0f;
    o.lin_s[i] = i < 7 ? y[i + 1] - y[i] : 0.0f;
  }
  return o;
}

int acquire(fecl_ldpc* h, std::unique_ptr<Workspace>* out) {
  {
    std::lock_guard<std::mutex> lk(h->mu);
    if (!h->pool.empty()) {
      *out = std::move(h->pool.back());
      h->pool.pop_back();
      return FECL_OK;
    }
  }
  std::unique_ptr<Workspace> ws(new Workspace());
  for (auto& e : ws->ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&ws->done, hipEventDisableTiming));
  HIP_TRY(hipEventRecord(ws->done, nullptr));
  *out = std::move(ws);
  return FECL_OK;
}

void release(fecl_ldpc* h, std::unique_ptr<Workspace> ws) {
  std::lock_guard<std::mutex> lk(h->mu);
  h->pool.push_back(std::move(ws));
}


// Device layout of the handle's current options.
Layout layout_of(const fecl_ldpc* h) {
  Layout lay;
  lay.fmt = h->msg_format;
  lay.words = h->words;
  lay.vpw = h->vpw;
  return lay;
}

namespace {

// The per-tile buffers of area a of a workspace (0: the first region's, 1: the spare).
struct AreaRefs {
  DevBuf *msg, *L, *hard, *phase, *iter, *fail, *flags, *out_hard, *out_ext, *out_app;
};
AreaRefs area_of(Workspace* ws, int a) {
  if (a == 0)
    return {&ws->msg, &ws->L, &ws->hard, &ws->phase, &ws->iter, &ws->fail, &ws->flags, &ws->out_hard, &ws->out_ext,
            &ws->out_app};
  Area& x = ws->spare;
  return {&x.msg, &x.L, &x.hard, &x.phase, &x.iter, &x.fail, &x.flags, &x.out_hard, &x.out_ext, &x.out_app};
}

// Sizes an area for `tiles` tiles.
int size_area(const Code& c, const Layout& lay, const AreaRefs& b, int64_t tiles, bool soft) {
  const int FPL = lay.fpl(), TW = lay.tw();
  HIP_TRY(b.msg->ensure((size_t)tiles * c.E * lay.rw() * 4));
  HIP_TRY(b.L->ensure((size_t)tiles * c.N * TW * 4));
  HIP_TRY(b.hard->ensure((size_t)tiles * c.N * FPL * 8));
  HIP_TRY(b.phase->ensure((size_t)tiles * TW));
  HIP_TRY(b.iter->ensure((size_t)tiles * TW * 4));
  HIP_TRY(b.fail->ensure((size_t)tiles * FPL * 8));
  HIP_TRY(b.flags->ensure((size_t)tiles * 8));
  HIP_TRY(b.out_hard->ensure((size_t)tiles * c.K * FPL * 8));
  if (soft) {
    HIP_TRY(b.out_ext->ensure((size_t)tiles * c.N * TW * 4));
    HIP_TRY(b.out_app->ensure((size_t)tiles * c.K * TW * 4));
  }
  return FECL_OK;
}

// Device views of a region of `tiles` tiles at the start of an area sized for at least that many.
void view_area(const AreaRefs& b, int64_t tiles, int32_t* active, bool soft, DevState* sp, DevOut* op) {
  DevState& s = *sp;
  s = DevState{};
  s.msg = b.msg->as<uint32_t>();
  s.L = b.L->as<float>();
  s.hard = b.hard->as<uint64_t>();
  s.phase = b.phase->as<uint8_t>();
  s.iter = b.iter->as<int32_t>();
  s.fail = b.fail->as<uint64_t>();
  s.tile_flags = b.flags->as<uint32_t>();
  s.tile_iter = b.flags->as<int32_t>() + tiles;
  s.active = active;
  s.tiles = (int32_t)tiles;
  DevOut& out = *op;
  out = DevOut{};
  out.hard = b.out_hard->as<uint64_t>();
  out.ext = soft ? b.out_ext->as<float>() : nullptr;
  out.app = soft ? b.out_app->as<float>() : nullptr;
}

}  // namespace

// Tiles of the largest pass run_passes makes of `frames` frames.
int64_t pass_tiles(const fecl_ldpc* h, const Layout& lay, int64_t frames) {
  const int64_t tw = lay.tw();
  int64_t pass = h->pass_frames > 0 ? h->pass_frames : kMaxPassFrames;
  pass = std::max<int64_t>(tw, (pass + tw - 1) / tw * tw);
  return (std::min(pass, frames) + tw - 1) / tw;
}

int prepare_graph_compaction(fecl_ldpc* h, Workspace* ws, int64_t frames) {
  const Layout lay = layout_of(h);
  const int64_t tiles = pass_tiles(h, lay, frames);
  for (DevBuf* b : {&ws->map[0], &ws->map[1], &ws->src_of}) HIP_TRY(b->ensure((size_t)tiles * lay.tw() * 4));
  HIP_TRY(ws->ctl.ensure(sizeof(CompactCtl)));
  return FECL_OK;
}

// Sizes the workspace for `frames` frames and returns the device views of it.
int prepare(fecl_ldpc* h, Workspace* ws, const Layout& lay, int64_t frames, bool soft, DevState* sp, DevOut* op) {
  const int64_t tiles = (frames + lay.tw() - 1) / lay.tw();
  const AreaRefs a = area_of(ws, 0);
  if (int rc = size_area(h->code, lay, a, tiles, soft)) return rc;
  HIP_TRY(ws->active.ensure((size_t)h->iterations * 4));
  view_area(a, tiles, ws->active.as<int32_t>(), soft, sp, op);
  return FECL_OK;
}

// Decodes frames [frame0, frame0 + frames) of io on stream st with workspace ws.
// capture: the launches are being recorded into a hipGraph, so every step is enqueued (frames
// that finish early make the remaining steps cheap no-ops) and nothing waits on the host.
int run_batch(fecl_ldpc* h, Workspace* ws, hipStream_t st, const BatchIO& io, int64_t frames, int64_t frame0,
              bool capture, const DevOpts* base) {
  // A small eager batch with the automatic layout takes fewer words per lane when that cuts its
  // padded slots (whole tiles) by more than 10 %: every iteration moves whole tiles, so 16
  // frames in a 256-frame tile would move 16 times the bytes they need.  Results do not depend
  // on the layout.  Pinned layouts (set_layout, autotune) and captured graphs, whose
  // workspaces are sized before capture, keep the handle's layout.
  Layout lay = layout_of(h);
  if (!capture && !h->words_pinned) {
    auto padded = [&](int w) { const int64_t tw = 64LL * w * lay.pw(); return (frames + tw - 1) / tw * tw; };
    while (lay.words > 1 && padded(lay.words / 2) * 11 < padded(lay.words) * 10) lay.words /= 2;
    if (!vpw_fits(lay, h->code.max_var_degree)) lay.vpw = 1;
  }
  const Code& c = h->code;
  const bool soft = io.res.msg || io.res.syst || io.res.parity;
  DevState s;
  DevOut out;
  if (int rc = prepare(h, ws, lay, frames, soft, &s, &out)) return rc;
  // one entry per tile per step: regions only shrink, so the first region's tiles bound them
  const int64_t live_entries = (int64_t)h->iterations * s.tiles;
  if (!capture && ws->host_live_cap < live_entries) {
    // growing: control kernels still queued from an earlier decode may write the old buffer
    if (ws->host_live) HIP_TRY(hipDeviceSynchronize());
    if (ws->host_live) (void)hipHostFree(ws->host_live);
    ws->host_live = ws->host_live_dev = nullptr;
    ws->host_live_cap = 0;
    HIP_TRY(hipHostMalloc((void**)&ws->host_live, (size_t)live_entries * 4, hipHostMallocMapped));
    HIP_TRY(hipHostGetDevicePointer((void**)&ws->host_live_dev, ws->host_live, 0));
    ws->host_live_cap = live_entries;
  }
  DevOpts o = base ? *base : make_opts(h);
  o.nt_rows = nt_rows_for(h, lay, s.tiles);
  const bool prof = h->profiling && !capture;
  if (prof) {
    while ((int)ws->pev.size() < h->iterations) {
      std::array<hipEvent_t, 4> ev{};
      for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
      ws->pev.push_back(ev);
    }
  }
  int steps_run = 0;
  // eager decodes: where each step's per-tile live counts sit in ws->host_live, and how many
  std::vector<int64_t> step_off(capture ? 0 : (size_t)h->iterations);
  std::vector<int32_t> step_tiles(capture ? 0 : (size_t)h->iterations);
  int64_t live_off = 0;
  // Compaction (eager decodes): once the live frames two steps back fit in three quarters of
  // the current region's tiles or fewer, and the tile-steps this frees outweigh moving them,
  // they move into the other area and later steps run on that region only.  The region left
  // behind hands its finished frames' results to the caller first.  Regions shrink by a
  // quarter or more each time, so area 1 needs three quarters of the first region's tiles.
  const int TW = lay.tw();
  const bool compact = h->compaction && !capture && !o.fused_syn && s.tiles >= 2;
  if (compact) {
    const int64_t spare = (3 * (int64_t)s.tiles + 3) / 4;
    if (int rc = size_area(c, lay, area_of(ws, 1), spare, soft)) return rc;
    for (DevBuf* b : {&ws->map[0], &ws->map[1], &ws->src_of}) HIP_TRY(b->ensure((size_t)spare * TW * 4));
  }
  DevState cur = s;
  DevOut cur_out = out;
  int cur_area = 0;
  const int32_t* cur_map = nullptr;
  // Captured graphs: no host sees the live counts, so after every kGraphCompactEvery-th step
  // the device decides and compacts in place (live frames into the region's first slots, in
  // one pass); a checkpoint that declines costs three launches that return at once.
  const size_t slot_bytes = (size_t)s.tiles * TW * 4;
  const bool graph_compact = h->compaction && capture && !o.fused_syn && s.tiles >= 2 &&
                             ws->map[0].bytes >= slot_bytes && ws->map[1].bytes >= slot_bytes &&
                             ws->src_of.bytes >= slot_bytes && ws->ctl.bytes >= sizeof(CompactCtl);
  int checkpoint = 0;

  HIP_TRY(kernels_fill_i32(s.active, h->iterations, 0, st));
  uint8_t init = (uint8_t)((h->iterations == 1 ? PH_FINAL : PH_RUN) | (io.state ? PH_VNINIT : PH_FRESH));
  HIP_TRY(kernels_ingest(lay, h->dc, s, io.llr, io.syst, io.dbl_in, frames, frame0, init, st));
  if (io.state) {
    HIP_TRY(kernels_state_in(lay, h->dc, s, io.state, io.dbl_in, frames, frame0, st));
    HIP_TRY(kernels_vn_init(lay, h->dc, s, c.max_var_degree, st));
  }
  for (int step = 0; step < h->iterations; ++step) {
    if (prof) HIP_TRY(hipEventRecord(ws->pev[step][0], st));
    HIP_TRY(kernels_check(lay, h->dc, cur, o, c.max_row_degree, st));
    if (prof) HIP_TRY(hipEventRecord(ws->pev[step][1], st));
    HIP_TRY(kernels_variable(lay, h->dc, cur, cur_out, c.max_var_degree, o.fused_syn, (o.nt_rows & 2) != 0, st));
    if (prof) HIP_TRY(hipEventRecord(ws->pev[step][2], st));
    DevState cs = cur;
    if (!capture) {   // this step's live frames per tile, for the host's count
      step_off[step] = live_off;
      step_tiles[step] = cur.tiles;
      cs.live_out = ws->host_live_dev + live_off;
      live_off += cur.tiles;
    }
    HIP_TRY(kernels_syndrome_control(lay, h->dc, cs, h->iterations, o.fused_syn, step, c.max_row_degree, st));
    if (prof) HIP_TRY(hipEventRecord(ws->pev[step][3], st));
    steps_run = step + 1;
    if (capture) {
      if (graph_compact && (step + 1) % kGraphCompactEvery == 0 && step + kCompactMinSteps < h->iterations) {
        int32_t* next_map = ws->map[checkpoint & 1].as<int32_t>();
        CompactCtl* ctl = ws->ctl.as<CompactCtl>();
        int32_t* src_of = ws->src_of.as<int32_t>();
        HIP_TRY(kernels_compact_plan(lay, cur, cur_map, next_map, src_of, ctl, h->iterations - 1 - step, st));
        HIP_TRY(kernels_results(lay, h->dc, cur, cur_out, frames, frame0, io.res, io.dbl_out, cur_map, ctl, st));
        HIP_TRY(kernels_compact_in_place(lay, h->dc, cur, src_of, ctl, st));
        cur_map = next_map;
        ++checkpoint;
      }
      continue;
    }
    HIP_TRY(hipEventRecord(ws->ev[step % 4], st));
    // Stop enqueuing once every frame of the batch has finished: the count of step-2 is read
    // while the last two steps run, so the device queue never drains on the host.
    if (step >= 2) {
      HIP_TRY(hipEventSynchronize(ws->ev[(step - 2) % 4]));
      int64_t live = 0;
      for (int32_t t = 0; t < step_tiles[step - 2]; ++t) live += ws->host_live[step_off[step - 2] + t];
      if (live == 0) break;
      // Live counts only fall, so the frames live now fit in `need` tiles.  Moving a frame
      // costs about one frame-iteration of memory traffic; it is counted 1.5 times.
      const int64_t need = (live + TW - 1) / TW;
      const int64_t freed = cur.tiles - need, remaining = h->iterations - 1 - step;
      if (compact && need <= 3 * (int64_t)cur.tiles / 4 && 2 * freed * remaining * TW > 3 * live &&
          step + kCompactMinSteps < h->iterations) {
        const int next = 1 - cur_area;
        DevState to;
        DevOut to_out;
        view_area(area_of(ws, next), need, s.active, soft, &to, &to_out);
        int32_t* next_map = ws->map[next].as<int32_t>();
        HIP_TRY(kernels_compact(lay, h->dc, cur, to, cur_map, next_map, ws->src_of.as<int32_t>(), st));
        HIP_TRY(kernels_results(lay, h->dc, cur, cur_out, frames, frame0, io.res, io.dbl_out, cur_map, nullptr, st));
        cur = to;
        cur_out = to_out;
        cur_area = next;
        cur_map = next_map;
        o.nt_rows = nt_rows_for(h, lay, cur.tiles);
      }
    }
  }
  HIP_TRY(kernels_results(lay, h->dc, cur, cur_out, frames, frame0, io.res, io.dbl_out, cur_map, nullptr, st));
  if (prof) {
    // kernel durations from events on the launch stream: check, variable, syndrome+control
    HIP_TRY(hipStreamSynchronize(st));
    std::lock_guard<std::mutex> lk(h->mu);
    for (int step = 0; step < steps_run; ++step) {
      for (int k = 0; k < 3; ++k) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ws->pev[step][k], ws->pev[step][k + 1]));
        h->prof_ms[k] += ms;
        h->prof_launches[k] += 1;
      }
    }
  }
  return FECL_OK;
}

int run_passes(fecl_ldpc* h, Workspace* ws, hipStream_t st, const BatchIO& io, int64_t frames, int64_t frame0,
               bool capture, const DevOpts* base) {
  const int64_t tw = layout_of(h).tw();
  int64_t pass = h->pass_frames > 0 ? h->pass_frames : kMaxPassFrames;
  pass = std::max<int64_t>(tw, (pass + tw - 1) / tw * tw);   // as pass_tiles
  for (int64_t f = 0; f < frames; f += pass)
    if (int rc = run_batch(h, ws, st, io, std::min(pass, frames - f), frame0 + f, capture, base)) return rc;
  return FECL_OK;
}

int copy_opts(const fecl_ldpc* h, OptsCopy* out) {
  out->opts = make_opts(h);
  const size_t n = h->sf.size();
  HIP_TRY(out->sf.ensure(n * sizeof(double)));
  HIP_TRY(out->sf32.ensure(n * sizeof(float)));
  HIP_TRY(hipMemcpy(out->sf.p, h->d_sf.p, n * sizeof(double), hipMemcpyDeviceToDevice));
  HIP_TRY(hipMemcpy(out->sf32.p, h->d_sf32.p, n * sizeof(float), hipMemcpyDeviceToDevice));
  out->opts.sf = out->sf.as<double>();
  out->opts.sf32 = out->sf32.as<float>();
  return FECL_OK;
}

int32_t nt_rows_for(const fecl_ldpc* h, const Layout& lay, int64_t tiles) {
  if (h->cache_policy >= 0) return h->cache_policy;
  const int64_t bytes = tiles * ((int64_t)h->code.E * lay.rw() * 4 + (int64_t)h->code.N * lay.tw() * 4);
  return bytes >= kNtBatchBytes ? 3 : 0;
}

int check_handle(const fecl_ldpc* h) {
  if (!h) return fail(FECL_INVALID_ARGUMENT, "null handle");
  return FECL_OK;
}

// Host-buffer decode/soDecode (FeCl's hooks hand over host FP64 buffers).  The frames go
// through the device in chunks, three stages deep:
//   in : host threads round chunk k + 1's FP64 inputs to FP32 (the device would round them the
//        same way) into pinned memory, and a copy stream moves them to device staging;
//   run: chunk k decodes, its outputs (FP32 values, bits, counts) into device staging;
//   out: a third stream moves chunk k - 1's outputs to pinned memory, and host threads widen
//        them to FP64 (exact) and copy them into the caller's buffers.
// Inputs and outputs cross PCIe as FP32 at the DMA rate of pinned memory; every stage has two
// buffers, used by chunks of alternating parity.
constexpr int64_t kHostChunkFrames = 512;
constexpr int64_t kHostFirstChunkFrames = 256;   // smaller: its conversion overlaps nothing
int host_decode(fecl_ldpc* h, const double* llr, const double* syst_in, const double* state_in, int64_t frames,
                uint8_t* msg_bits, double* msg, double* syst, double* parity, double* state, int32_t* iterations) {
  if (int rc = check_handle(h)) return rc;
  if (frames < 0) return fail(FECL_INVALID_ARGUMENT, "Invalid size for parity");
  if (frames == 0) return FECL_OK;
  if (!llr) return fail(FECL_INVALID_ARGUMENT, "Input must contains parity");
  const Code& c = h->code;
  HIP_TRY(hipSetDevice(h->device));
  struct InFlight {
    InFlight() { ++g_host_calls; }
    ~InFlight() { --g_host_calls; }
  } in_flight;
  std::unique_ptr<Workspace> ws;
  if (int rc = acquire(h, &ws)) return rc;
  struct Back { fecl_ldpc* h; std::unique_ptr<Workspace>* ws; ~Back() { release(h, std::move(*ws)); } } back{h, &ws};
  if (!ws->own_stream) HIP_TRY(hipStreamCreateWithFlags(&ws->own_stream, hipStreamNonBlocking));
  if (!ws->copy_stream) {
    HIP_TRY(hipStreamCreateWithFlags(&ws->copy_stream, hipStreamNonBlocking));
    for (int b = 0; b < 2; ++b) {
      HIP_TRY(hipEventCreateWithFlags(&ws->copied[b], hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&ws->used[b], hipEventDisableTiming));
    }
  }
  if (!ws->out_stream) {
    HIP_TRY(hipStreamCreateWithFlags(&ws->out_stream, hipStreamNonBlocking));
    for (int b = 0; b < 2; ++b) {
      HIP_TRY(hipEventCreateWithFlags(&ws->decoded[b], hipEventDisableTiming));
      HIP_TRY(hipEventCreateWithFlags(&ws->drained[b], hipEventDisableTiming));
    }
  }
  hipStream_t st = ws->own_stream, cs = ws->copy_stream, os = ws->out_stream;
  const int64_t tw = layout_of(h).tw();
  const int64_t chunk = std::min<int64_t>(frames, std::max<int64_t>(tw, kHostChunkFrames / tw * tw));
  const int64_t first = std::min<int64_t>(chunk, std::max<int64_t>(tw, kHostFirstChunkFrames / tw * tw));
  // frames [start(k), start(k) + size(k)) form chunk k
  auto start = [&](int64_t k) { return k == 0 ? 0 : first + (k - 1) * chunk; };
  auto size = [&](int64_t k) { return std::min(k == 0 ? first : chunk, frames - start(k)); };
  const int64_t nchunks = 1 + (frames - first + chunk - 1) / chunk;
  // output sections of one chunk (same offsets in device staging and in pinned memory)
  size_t out_bytes = 0;
  auto take = [&](bool want, size_t n) { size_t off = out_bytes; if (want) out_bytes += (n + 255) / 256 * 256; return off; };
  const size_t o_bits = take(msg_bits != nullptr, (size_t)chunk * c.K);
  const size_t o_iter = take(iterations != nullptr, (size_t)chunk * 4);
  const size_t o_msg = take(msg != nullptr, (size_t)chunk * c.K * 4);
  const size_t o_syst = take(syst != nullptr, (size_t)chunk * c.K * 4);
  const size_t o_par = take(parity != nullptr, (size_t)chunk * c.N * 4);
  const size_t o_state = take(state != nullptr, (size_t)chunk * c.E * 4);
  for (int b = 0; b < 2; ++b) {
    HIP_TRY(ws->stage_in[b].ensure((size_t)chunk * c.N * 4));
    HIP_TRY(ws->pin_in[b].ensure((size_t)chunk * c.N * 4));
    if (state_in) {
      HIP_TRY(ws->stage_state[b].ensure((size_t)chunk * c.E * 4));
      HIP_TRY(ws->pin_state[b].ensure((size_t)chunk * c.E * 4));
    }
    HIP_TRY(ws->stage_out[b].ensure(std::max<size_t>(out_bytes, 256)));
    HIP_TRY(ws->pin_out[b].ensure(std::max<size_t>(out_bytes, 256)));
  }
  // (errors are recorded per thread, so helpers hand their message back with the status)
  using Status = std::pair<int, std::string>;
  auto helper = [&](auto fn) {
    return std::async(std::launch::async, [&, fn]() -> Status {
      if (hipSetDevice(h->device) != hipSuccess) return {FECL_DEVICE_ERROR, "hipSetDevice failed"};
      const int rc = fn();
      return {rc, rc ? g_err : std::string()};
    });
  };
  // in: chunk k's inputs, FP32 into pinned buffer k & 1 once chunk k - 2's copy out of it is
  // done, then to device staging k & 1 once chunk k - 2's decode has read that
  auto copy_in = [&](int64_t k) {
    return helper([&, k]() -> int {
      const int b = (int)(k & 1);
      const int64_t f0 = start(k), n = size(k);
      if (k >= 2) HIP_TRY(hipEventSynchronize(ws->copied[b]));
      // the a-priori values are added here, in double, and the sum rounded once
      if (syst_in)
        to_float_plus(llr + f0 * c.N, syst_in + f0 * c.K, ws->pin_in[b].as<float>(), n, c.N, c.K);
      else
        to_float(llr + f0 * c.N, ws->pin_in[b].as<float>(), (size_t)n * c.N);
      if (state_in) to_float(state_in + f0 * c.E, ws->pin_state[b].as<float>(), (size_t)n * c.E);
      if (k >= 2) HIP_TRY(hipStreamWaitEvent(cs, ws->used[b], 0));
      HIP_TRY(hipMemcpyAsync(ws->stage_in[b].p, ws->pin_in[b].p, (size_t)n * c.N * 4, hipMemcpyHostToDevice, cs));
      if (state_in)
        HIP_TRY(hipMemcpyAsync(ws->stage_state[b].p, ws->pin_state[b].p, (size_t)n * c.E * 4, hipMemcpyHostToDevice, cs));
      HIP_TRY(hipEventRecord(ws->copied[b], cs));
      return FECL_OK;
    });
  };
  // out: chunk k's outputs from pinned buffer k & 1 (filled by the out stream) into the
  // caller's buffers
  auto copy_out = [&](int64_t k) {
    return helper([&, k]() -> int {
      const int b = (int)(k & 1);
      const int64_t f0 = start(k), n = size(k);
      HIP_TRY(hipEventSynchronize(ws->drained[b]));
      const char* pb = ws->pin_out[b].as<char>();
      if (msg_bits) copy_bytes(pb + o_bits, msg_bits + f0 * c.K, (size_t)n * c.K);
      if (iterations) std::memcpy(iterations + f0, pb + o_iter, (size_t)n * 4);
      if (msg) to_double((const float*)(pb + o_msg), msg + f0 * c.K, (size_t)n * c.K);
      if (syst) to_double((const float*)(pb + o_syst), syst + f0 * c.K, (size_t)n * c.K);
      if (parity) to_double((const float*)(pb + o_par), parity + f0 * c.N, (size_t)n * c.N);
      if (state) to_double((const float*)(pb + o_state), state + f0 * c.E, (size_t)n * c.E);
      return FECL_OK;
    });
  };
  std::future<Status> next_in = copy_in(0), out_done[2];
  // waits for every helper still running (they use the workspace and the caller's buffers)
  auto settle = [&]() {
    Status worst{FECL_OK, std::string()};
    for (std::future<Status>* f : {&next_in, &out_done[0], &out_done[1]})
      if (f->valid()) {
        Status r = f->get();
        if (r.first && !worst.first) worst = r;
      }
    return worst;
  };
  auto bail = [&](int rc) {
    const std::string msg_here = g_err;
    settle();
    (void)hipStreamSynchronize(cs);
    (void)hipStreamSynchronize(os);
    (void)hipStreamSynchronize(st);
    return fail(rc, msg_here);
  };
  for (int64_t k = 0; k < nchunks; ++k) {
    const int b = (int)(k & 1);
    const int64_t n = size(k);
    const Status got = next_in.get();
    if (got.first) {
      g_err = got.second;
      return bail(got.first);
    }
    if (k + 1 < nchunks) next_in = copy_in(k + 1);   // overlaps this chunk's decode
    if (hipStreamWaitEvent(st, ws->copied[b], 0) != hipSuccess ||
        (k >= 2 && hipStreamWaitEvent(st, ws->drained[b], 0) != hipSuccess)) {
      g_err = "hipStreamWaitEvent failed";
      return bail(FECL_DEVICE_ERROR);
    }
    char* ob = ws->stage_out[b].as<char>();
    BatchIO io{};
    io.llr = ws->stage_in[b].p;
    io.syst = nullptr;   // already added to the channel values on the host
    io.state = state_in ? ws->stage_state[b].p : nullptr;
    io.dbl_in = false;
    io.dbl_out = false;
    io.res.msg_bits = msg_bits ? (uint8_t*)(ob + o_bits) : nullptr;
    io.res.iterations = iterations ? (int32_t*)(ob + o_iter) : nullptr;
    io.res.msg = msg ? (void*)(ob + o_msg) : nullptr;
    io.res.syst = syst ? (void*)(ob + o_syst) : nullptr;
    io.res.parity = parity ? (void*)(ob + o_par) : nullptr;
    io.res.state = state ? (void*)(ob + o_state) : nullptr;
    if (int rc = run_passes(h, ws.get(), st, io, n, 0, false)) return bail(rc);
    if (hipEventRecord(ws->used[b], st) != hipSuccess || hipEventRecord(ws->decoded[b], st) != hipSuccess) {
      g_err = "hipEventRecord failed";
      return bail(FECL_DEVICE_ERROR);
    }
    // pinned buffer b is free once chunk k - 2's outputs have been copied out of it
    if (out_done[b].valid()) {
      const Status r = out_done[b].get();
      if (r.first) {
        g_err = r.second;
        return bail(r.first);
      }
    }
    char* pb = ws->pin_out[b].as<char>();
    bool ok = hipStreamWaitEvent(os, ws->decoded[b], 0) == hipSuccess;
    auto d2h = [&](bool want, size_t off, size_t bytes) {
      if (want && ok) ok = hipMemcpyAsync(pb + off, ob + off, bytes, hipMemcpyDeviceToHost, os) == hipSuccess;
    };
    d2h(msg_bits != nullptr, o_bits, (size_t)n * c.K);
    d2h(iterations != nullptr, o_iter, (size_t)n * 4);
    d2h(msg != nullptr, o_msg, (size_t)n * c.K * 4);
    d2h(syst != nullptr, o_syst, (size_t)n * c.K * 4);
    d2h(parity != nullptr, o_par, (size_t)n * c.N * 4);
    d2h(state != nullptr, o_state, (size_t)n * c.E * 4);
    if (!ok || hipEventRecord(ws->drained[b], os) != hipSuccess) {
      g_err = "device-to-host copy failed";
      return bail(FECL_DEVICE_ERROR);
    }
    out_done[b] = copy_out(k);
  }
  const Status r = settle();
  if (r.first) return fail(r.first, r.second);
  HIP_TRY(hipStreamSynchronize(cs));
  HIP_TRY(hipStreamSynchronize(os));
  HIP_TRY(hipStreamSynchronize(st));
  return FECL_OK;
}

int device_decode(fecl_ldpc* h, const float* llr, const float* syst_in, const float* state_in, int64_t frames,
                  uint8_t* msg_bits, float* msg, float* syst, float* parity, float* state, int32_t* iterations,
                  void* stream) {
  if (int rc = check_handle(h)) return rc;
  if (frames < 0) return fail(FECL_INVALID_ARGUMENT, "Invalid size for parity");
  if (frames == 0) return FECL_OK;
  if (!llr) return fail(FECL_INVALID_ARGUMENT, "Input must contains parity");
  HIP_TRY(hipSetDevice(h->device));
  std::unique_ptr<Workspace> ws;
  if (int rc = acquire(h, &ws)) return rc;
  struct Back { fecl_ldpc* h; std::unique_ptr<Workspace>* ws; ~Back() { release(h, std::move(*ws)); } } back{h, &ws};
  hipStream_t st = (hipStream_t)stream;
  HIP_TRY(hipStreamWaitEvent(st, ws->done, 0));
  BatchIO io{};
  io.llr = llr;
  io.syst = syst_in;
  io.state = state_in;
  io.dbl_in = false;
  io.dbl_out = false;
  io.res = BatchResults{msg_bits, iterations, msg, syst, parity, state};
  if (int rc = run_passes(h, ws.get(), st, io, frames, 0, false)) return rc;
  HIP_TRY(hipEventRecord(ws->done, st));
  return FECL_OK;
}

}  // namespace engine
}  // namespace fecl

using namespace fecl;
using namespace fecl::engine;

extern "C" {

const char* fecl_last_error(void) { return g_err.c_str(); }

int fecl_dvbs2_matrix(int32_t n, int32_t rate_num, int32_t rate_den, int64_t* dims, int32_t* row_ptr, int32_t* col) {
  if (!dims) return fail(FECL_INVALID_ARGUMENT, "null dims");
  if (!dvbs2_matrix(n, rate_num, rate_den, dims, row_ptr, col)) {
    if (n != 64800 && n != 16200) return fail(FECL_INVALID_ARGUMENT, "Invalid length");
    return fail(FECL_INVALID_ARGUMENT, "Invalid rate");
  }
  return FECL_OK;
}

int fecl_gallager_matrix(int64_t n, int64_t wc, int64_t wr, uint64_t seed, int64_t* dims, int32_t* row_ptr,
                         int32_t* col) {
  if (!dims) return fail(FECL_INVALID_ARGUMENT, "null dims");
  Csr m;
  std::string err = gallager_matrix(n, wc, wr, seed, &m);
  if (!err.empty()) return fail(FECL_INVALID_ARGUMENT, err);
  dims[0] = m.rows();
  dims[1] = n;
  dims[2] = (int64_t)m.idx.size();
  if (row_ptr && col) {
    std::copy(m.ptr.begin(), m.ptr.end(), row_ptr);
    std::copy(m.idx.begin(), m.idx.end(), col);
  }
  return FECL_OK;
}

int fecl_ldpc_structure(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, int64_t* dims,
                        int32_t* out_row_ptr, int32_t* out_col) {
  if (!dims) return fail(FECL_INVALID_ARGUMENT, "null dims");
  GeneratorForm g;
  std::string err = generator_form(row_ptr, col, M, N, &g);
  if (!err.empty()) return fail(FECL_INVALID_ARGUMENT, err);
  dims[0] = g.M;
  dims[1] = g.N;
  dims[2] = g.K;
  dims[3] = (int64_t)g.H.idx.size();
  if (out_row_ptr && out_col) {
    std::copy(g.H.ptr.begin(), g.H.ptr.end(), out_row_ptr);
    std::copy(g.H.idx.begin(), g.H.idx.end(), out_col);
  }
  return FECL_OK;
}

int fecl_encode(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, const uint8_t* msg,
                int64_t frames, uint8_t* parity) {
  GeneratorForm g;
  std::string err = generator_form(row_ptr, col, M, N, &g);
  if (!err.empty()) return fail(FECL_INVALID_ARGUMENT, err);
  if (frames < 0 || (frames > 0 && (!msg || !parity))) return fail(FECL_INVALID_ARGUMENT, "Invalid size for message");
  for (int64_t f = 0; f < frames; ++f) encode_generator(g, msg + f * g.K, parity + f * g.N);
  return FECL_OK;
}

int fecl_ldpc_scaling_table(const int32_t* row_ptr, int64_t M, int64_t iterations, const int64_t* keys,
                            const int64_t* lengths, const int64_t* offsets, const double* values, int64_t entries,
                            double* table, int64_t table_capacity, int32_t* rows, int32_t* cols) {
  if (!row_ptr || M < 1 || (entries > 0 && (!keys || !lengths || !offsets || !values)) || !rows || !cols)
    return fail(FECL_INVALID_ARGUMENT, "invalid arguments");
  Code c;
  c.M = M;
  c.row_ptr.assign(row_ptr, row_ptr + M + 1);
  for (int64_t r = 0; r < M; ++r) c.max_row_degree = std::max(c.max_row_degree, row_ptr[r + 1] - row_ptr[r]);
  if (c.max_row_degree < 2) return fail(FECL_INVALID_ARGUMENT, "every check row needs at least two edges");
  std::unordered_map<size_t, std::vector<double>> map;
  for (int64_t i = 0; i < entries; ++i)
    map[(size_t)keys[i]] = std::vector<double>(values + offsets[i], values + offsets[i] + lengths[i]);
  std::vector<double> t;
  std::string err = scaling_map_to_table(map, c, iterations, &t, rows, cols);
  if (!err.empty()) return fail(FECL_INVALID_ARGUMENT, err);
  if (table) {
    if ((int64_t)t.size() > table_capacity) return fail(FECL_INVALID_ARGUMENT, "scaling table buffer too small");
    std::copy(t.begin(), t.end(), table);
  }
  return FECL_OK;
}

}  // extern "C"

namespace {

// fec::Ldpc(const Structure&) (src/Ldpc.h:83): a decoder for a structure already in generator
// form (from fecl_ldpc_create's transform or from a saved codec).
int create_from_form(GeneratorForm&& form, const fecl_decoder_options* options, fecl_ldpc** out) {
  if (!out) return fail(FECL_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  std::unique_ptr<fecl_ldpc> h(new fecl_ldpc());
  h->gen = std::move(form);
  std::string err = build_code(h->gen.H.ptr.data(), h->gen.H.idx.data(), h->gen.M, h->gen.N, &h->code, h->gen.K);
  if (!err.empty()) return fail(FECL_INVALID_ARGUMENT, err);
  h->gen.H = Csr();  // the code holds it
  const Code& c = h->code;
  if (c.max_var_degree > 32) return fail(FECL_NOT_SUPPORTED, "variable degree above 32 is not supported");
  HIP_TRY(hipGetDevice(&h->device));
  HIP_TRY(h->d_row_ptr.ensure(c.row_ptr.size() * 4));
  HIP_TRY(h->d_col.ensure(c.col.size() * 4));
  HIP_TRY(h->d_var_ptr.ensure(c.var_ptr.size() * 4));
  HIP_TRY(h->d_var_edges.ensure(c.var_edges.size() * 4));
  HIP_TRY(hipMemcpy(h->d_row_ptr.p, c.row_ptr.data(), c.row_ptr.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_col.p, c.col.data(), c.col.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_var_ptr.p, c.var_ptr.data(), c.var_ptr.size() * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_var_edges.p, c.var_edges.data(), c.var_edges.size() * 4, hipMemcpyHostToDevice));
  h->dc.row_ptr = h->d_row_ptr.as<int32_t>();
  h->dc.col = h->d_col.as<int32_t>();
  h->dc.var_ptr = h->d_var_ptr.as<int32_t>();
  h->dc.var_edges = h->d_var_edges.as<int32_t>();
  {
    const int vs = var_table_stride(c.max_var_degree);
    // padded by a few variables so a wavefront serving several variables reads past none
    std::vector<int32_t> tab((size_t)(c.N + 4) * vs, -1);
    for (int64_t v = 0; v < c.N; ++v)
      for (int32_t j = c.var_ptr[v]; j < c.var_ptr[v + 1]; ++j) tab[(size_t)v * vs + (j - c.var_ptr[v])] = c.var_edges[j];
    HIP_TRY(h->d_var_tab.ensure(tab.size() * 4));
    HIP_TRY(hipMemcpy(h->d_var_tab.p, tab.data(), tab.size() * 4, hipMemcpyHostToDevice));
    h->dc.var_tab = h->d_var_tab.as<int32_t>();
    h->dc.vstride = vs;
  }
  h->dc.M = (int32_t)c.M;
  h->dc.N = (int32_t)c.N;
  h->dc.K = (int32_t)c.K;
  h->dc.E = (int32_t)c.E;
  fecl_decoder_options def{FECL_APPROXIMATE, 1, nullptr, 1, c.max_row_degree - 1, FECL_MSG_F32};
  if (int rc = set_options(h.get(), options ? options : &def)) return rc;
  *out = h.release();
  return FECL_OK;
}

// The archive fields of a codec with options o on a code of maximum row degree maxdeg.
int archive_options(const fecl_decoder_options* o, int32_t maxdeg, int32_t wgs, ArchiveOptions* a) {
  if (!o) return fail(FECL_INVALID_ARGUMENT, "null options");
  if (o->algorithm < 0 || o->algorithm > 2) return fail(FECL_INVALID_ARGUMENT, "Invalid decoder algorithm");
  if (o->iterations < 1) return fail(FECL_INVALID_ARGUMENT, "iterations must be >= 1");
  a->algorithm = o->algorithm;
  a->iterations = o->iterations;
  a->cols = maxdeg - 1;
  if (o->scaling) {
    if (o->scaling_cols != a->cols || (o->scaling_rows != 1 && o->scaling_rows != o->iterations))
      return fail(FECL_INVALID_ARGUMENT, "Wrong size for scaling factor");
    a->rows = o->scaling_rows;
    a->scaling.assign(o->scaling, o->scaling + (size_t)(a->rows * a->cols));
  } else {
    a->rows = 1;
    a->scaling.assign((size_t)a->cols, 1.0);
  }
  a->work_group_size = wgs;
  return FECL_OK;
}

int put_archive(const std::vector<uint8_t>& bytes, uint8_t* buffer, int64_t capacity, int64_t* size) {
  if (!size) return fail(FECL_INVALID_ARGUMENT, "null size");
  *size = (int64_t)bytes.size();
  if (!buffer) return FECL_OK;
  if (capacity < (int64_t)bytes.size()) return fail(FECL_INVALID_ARGUMENT, "archive buffer too small");
  std::memcpy(buffer, bytes.data(), bytes.size());
  return FECL_OK;
}

}  // namespace

extern "C" {

int fecl_ldpc_create(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N,
                     const fecl_decoder_options* options, fecl_ldpc** out) {
  if (!out) return fail(FECL_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  // FeCl decodes and encodes with H in its generator form (src/detail/Ldpc.cpp:46-55); for
  // staircase codes such as DVB-S2 that form is H itself.
  GeneratorForm form;
  std::string err = generator_form(row_ptr, col, M, N, &form);
  if (!err.empty()) return fail(FECL_INVALID_ARGUMENT, err);
  return create_from_form(std::move(form), options, out);
}

int fecl_ldpc_create_checks(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, int64_t K,
                            const fecl_decoder_options* options, fecl_ldpc** out) {
  if (!out) return fail(FECL_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  GeneratorForm form;
  // the encoder blocks are those of the matrix's own form when that form is the matrix itself
  // (every staircase code); otherwise the handle decodes only
  std::string err = generator_form(row_ptr, col, M, N, &form);
  const bool own = err.empty() && form.K == K &&
                   std::equal(form.H.ptr.begin(), form.H.ptr.end(), row_ptr) &&
                   std::equal(form.H.idx.begin(), form.H.idx.end(), col);
  if (!own) {
    form = GeneratorForm();
    form.M = M;
    form.N = N;
    form.K = K;
    if (!row_ptr || !col || M < 1 || N <= M) return fail(FECL_INVALID_ARGUMENT, "parity-check matrix must have 1 <= rows < cols");
    form.H.ptr.assign(row_ptr, row_ptr + M + 1);
    form.H.idx.assign(col, col + row_ptr[M]);
  }
  const int rc = create_from_form(std::move(form), options, out);
  if (rc == FECL_OK) (*out)->has_encoder = own;
  return rc;
}

int fecl_ldpc_save(const fecl_ldpc* h, int32_t work_group_size, uint8_t* buffer, int64_t capacity, int64_t* size) {
  if (int rc = check_handle(h)) return rc;
  if (!h->has_encoder) return fail(FECL_NOT_SUPPORTED, "this handle was made without FeCl's encoder blocks");
  const Code& c = h->code;
  GeneratorForm g = h->gen;
  g.H.ptr = c.row_ptr;
  g.H.idx = c.col;
  ArchiveOptions a;
  a.algorithm = h->algorithm;
  a.iterations = h->iterations;
  a.rows = h->sf_rows;
  a.cols = h->sf_cols;
  a.scaling = h->sf;
  a.work_group_size = work_group_size;
  std::vector<uint8_t> bytes;
  archive_write(g, c.E, a, &bytes);
  return put_archive(bytes, buffer, capacity, size);
}

int fecl_ldpc_load(const uint8_t* buffer, int64_t size, fecl_ldpc** out, int32_t* work_group_size) {
  if (!out) return fail(FECL_INVALID_ARGUMENT, "null output handle");
  *out = nullptr;
  GeneratorForm g;
  ArchiveOptions a;
  std::string err = archive_read(buffer, size, &g, &a);
  if (!err.empty()) return fail(FECL_INVALID_ARGUMENT, err);
  if (a.iterations < 1 || a.iterations > INT32_MAX) return fail(FECL_INVALID_ARGUMENT, "iterations must be >= 1");
  fecl_decoder_options o{a.algorithm, (int32_t)a.iterations, a.scaling.data(), (int32_t)a.rows, (int32_t)a.cols,
                         FECL_MSG_F32};
  if (int rc = create_from_form(std::move(g), &o, out)) return rc;
  if (work_group_size) *work_group_size = a.work_group_size;
  return FECL_OK;
}

int fecl_archive_save(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N,
                      const fecl_decoder_options* options, int32_t work_group_size, uint8_t* buffer,
                      int64_t capacity, int64_t* size) {
  GeneratorForm g;
  std::string err = generator_form(row_ptr, col, M, N, &g);
  if (!err.empty()) return fail(FECL_INVALID_ARGUMENT, err);
  int32_t maxdeg = 0;
  for (int64_t r = 0; r < M; ++r) maxdeg = std::max(maxdeg, row_ptr[r + 1] - row_ptr[r]);
  ArchiveOptions a;
  if (int rc = archive_options(options, maxdeg, work_group_size, &a)) return rc;
  std::vector<uint8_t> bytes;
  archive_write(g, row_ptr[M], a, &bytes);
  return put_archive(bytes, buffer, capacity, size);
}

int fecl_archive_load(const uint8_t* buffer, int64_t size, int64_t* dims, int32_t* row_ptr, int32_t* col,
                      fecl_decoder_options* options, double* scaling, int64_t capacity, int32_t* work_group_size) {
  if (!dims) return fail(FECL_INVALID_ARGUMENT, "null dims");
  GeneratorForm g;
  ArchiveOptions a;
  std::string err = archive_read(buffer, size, &g, &a);
  if (!err.empty()) return fail(FECL_INVALID_ARGUMENT, err);
  dims[0] = g.M;
  dims[1] = g.N;
  dims[2] = g.K;
  dims[3] = (int64_t)g.H.idx.size();
  if (row_ptr && col) {
    std::copy(g.H.ptr.begin(), g.H.ptr.end(), row_ptr);
    std::copy(g.H.idx.begin(), g.H.idx.end(), col);
  }
  if (options) {
    options->algorithm = a.algorithm;
    options->iterations = (int32_t)std::min<int64_t>(a.iterations, INT32_MAX);
    options->scaling_rows = (int32_t)a.rows;
    options->scaling_cols = (int32_t)a.cols;
    options->msg_format = FECL_MSG_F32;
    options->scaling = nullptr;
    if (scaling) {
      if ((int64_t)a.scaling.size() > capacity) return fail(FECL_INVALID_ARGUMENT, "scaling buffer too small");
      std::copy(a.scaling.begin(), a.scaling.end(), scaling);
      options->scaling = scaling;
    }
  }
  if (work_group_size) *work_group_size = a.work_group_size;
  return FECL_OK;
}

int fecl_ldpc_destroy(fecl_ldpc* h) {
  if (h) {
    (void)hipSetDevice(h->device);
    delete h;
  }
  return FECL_OK;
}

int fecl_ldpc_set_options(fecl_ldpc* h, const fecl_decoder_options* options) {
  if (int rc = check_handle(h)) return rc;
  return set_options(h, options);
}

int fecl_ldpc_get_options(const fecl_ldpc* h, fecl_decoder_options* options, double* scaling, int64_t capacity) {
  if (int rc = check_handle(h)) return rc;
  if (!options) return fail(FECL_INVALID_ARGUMENT, "null options");
  options->algorithm = h->algorithm;
  options->iterations = h->iterations;
  options->scaling_rows = h->sf_rows;
  options->scaling_cols = h->sf_cols;
  options->msg_format = h->msg_format;
  options->scaling = nullptr;
  if (scaling) {
    if ((int64_t)h->sf.size() > capacity) return fail(FECL_INVALID_ARGUMENT, "scaling buffer too small");
    std::copy(h->sf.begin(), h->sf.end(), scaling);
    options->scaling = scaling;
  }
  return FECL_OK;
}

int fecl_ldpc_sizes(const fecl_ldpc* h, int64_t* msg, int64_t* syst, int64_t* parity, int64_t* state) {
  if (int rc = check_handle(h)) return rc;
  if (msg) *msg = h->code.K;
  if (syst) *syst = h->code.K;
  if (parity) *parity = h->code.N;
  if (state) *state = h->code.E;
  return FECL_OK;
}

int fecl_ldpc_checks(const fecl_ldpc* h, int64_t* dims, int32_t* row_ptr, int32_t* col) {
  if (int rc = check_handle(h)) return rc;
  if (!dims) return fail(FECL_INVALID_ARGUMENT, "null dims");
  const Code& c = h->code;
  dims[0] = c.M;
  dims[1] = c.N;
  dims[2] = c.K;
  dims[3] = c.E;
  if (row_ptr && col) {
    std::copy(c.row_ptr.begin(), c.row_ptr.end(), row_ptr);
    std::copy(c.col.begin(), c.col.end(), col);
  }
  return FECL_OK;
}

int fecl_ldpc_encode(const fecl_ldpc* h, const uint8_t* msg, int64_t frames, uint8_t* parity) {
  if (int rc = check_handle(h)) return rc;
  if (!h->has_encoder) return fail(FECL_NOT_SUPPORTED, "this handle was made without FeCl's encoder blocks");
  if (frames < 0 || (frames > 0 && (!msg || !parity))) return fail(FECL_INVALID_ARGUMENT, "Invalid size for message");
  for (int64_t f = 0; f < frames; ++f) encode_generator(h->gen, msg + f * h->code.K, parity + f * h->code.N);
  return FECL_OK;
}

int fecl_ldpc_encode_device(fecl_ldpc* h, const uint8_t* msg, int64_t frames, uint8_t* parity, void* stream) {
  if (int rc = check_handle(h)) return rc;
  if (frames < 0 || (frames > 0 && (!msg || !parity))) return fail(FECL_INVALID_ARGUMENT, "Invalid size for message");
  if (!h->code.staircase) return fail(FECL_NOT_SUPPORTED, "encoder supports staircase (DVB-S2 type) codes only");
  if (h->code.N > 65536) return fail(FECL_NOT_SUPPORTED, "device encoder holds a frame in 64 KiB of LDS");
  if (frames == 0) return FECL_OK;
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(kernels_encode(h->dc, msg, parity, frames, (hipStream_t)stream));
  return FECL_OK;
}

int fecl_ldpc_puncturing(int64_t syst_size, int64_t parity_size, const uint8_t* mask, int64_t mask_len,
                         const uint8_t* syst_mask, int64_t syst_mask_len, int64_t* sequence, int64_t capacity,
                         int64_t* count) {
  if (!count || syst_size < 0 || parity_size < syst_size || mask_len < 0 || syst_mask_len < 0 ||
      (mask_len > 0 && !mask) || (syst_mask_len > 0 && !syst_mask))
    return fail(FECL_INVALID_ARGUMENT, "invalid arguments");
  // FeCl's rule, index for index: systematic positions advance the index only when kept (and
  // the systematic mask is read at that index); parity positions always advance it.
  std::vector<int64_t> seq;
  int64_t idx = 0;
  for (int64_t i = 0; i < syst_size; ++i) {
    const bool keep = syst_mask_len == 0 ? (mask_len == 0 || mask[i % mask_len] != 0)
                                         : syst_mask[idx % syst_mask_len] != 0;
    if (keep) seq.push_back(idx++);
  }
  for (int64_t i = 0; i < parity_size - syst_size; ++i) {
    const bool keep = syst_mask_len == 0 ? (mask_len == 0 || mask[i % mask_len] != 0)
                                         : (mask_len == 0 || mask[idx % mask_len] != 0);
    if (keep) seq.push_back(idx);
    ++idx;
  }
  *count = (int64_t)seq.size();
  if (sequence) {
    if (capacity < (int64_t)seq.size()) return fail(FECL_INVALID_ARGUMENT, "sequence buffer too small");
    std::copy(seq.begin(), seq.end(), sequence);
  }
  return FECL_OK;
}

int fecl_permute_device(const int64_t* sequence, int64_t output_size, int64_t input_size, const float* in,
                        float* out, int64_t frames, int32_t inverse, void* stream) {
  if (frames < 0 || output_size < 0 || input_size < 0 || (frames > 0 && output_size > 0 && (!sequence || !in || !out)))
    return fail(FECL_INVALID_ARGUMENT, "invalid arguments");
  if (frames == 0) return FECL_OK;
  HIP_TRY(kernels_permute(sequence, output_size, input_size, in, out, frames, inverse != 0, (hipStream_t)stream));
  return FECL_OK;
}

int fecl_ldpc_check(const fecl_ldpc* h, const uint8_t* parity, int64_t frames, int32_t* ok) {
  if (int rc = check_handle(h)) return rc;
  if (!ok || frames < 0 || (frames > 0 && !parity)) return fail(FECL_INVALID_ARGUMENT, "Invalid size for parity");
  const Code& c = h->code;
  *ok = 1;
  for (int64_t f = 0; f < frames && *ok; ++f) {
    const uint8_t* x = parity + f * c.N;
    for (int64_t r = 0; r < c.M; ++r) {
      uint8_t s = 0;
      for (int32_t e = c.row_ptr[r]; e < c.row_ptr[r + 1]; ++e) s ^= x[c.col[e]] & 1;
      if (s) { *ok = 0; break; }
    }
  }
  return FECL_OK;
}

int fecl_ldpc_decode(fecl_ldpc* h, const double* llr, int64_t frames, uint8_t* msg, int32_t* iterations) {
  return host_decode(h, llr, nullptr, nullptr, frames, msg, nullptr, nullptr, nullptr, nullptr, iterations);
}

int fecl_ldpc_sodecode(fecl_ldpc* h, const double* parity_in, const double* syst_in, const double* state_in,
                       int64_t frames, double* msg_out, double* syst_out, double* parity_out, double* state_out,
                       int32_t* iterations) {
  return host_decode(h, parity_in, syst_in, state_in, frames, nullptr, msg_out, syst_out, parity_out, state_out,
                     iterations);
}

int fecl_ldpc_decode_device(fecl_ldpc* h, const float* llr, int64_t frames, uint8_t* msg, int32_t* iterations,
                            void* stream) {
  return device_decode(h, llr, nullptr, nullptr, frames, msg, nullptr, nullptr, nullptr, nullptr, iterations, stream);
}

int fecl_ldpc_sodecode_device(fecl_ldpc* h, const float* parity_in, const float* syst_in, const float* state_in,
                              int64_t frames, float* msg_out, float* syst_out, float* parity_out, float* state_out,
                              int32_t* iterations, void* stream) {
  return device_decode(h, parity_in, syst_in, state_in, frames, nullptr, msg_out, syst_out, parity_out, state_out,
                       iterations, stream);
}

int fecl_ldpc_profile(fecl_ldpc* h, int32_t enable) {
  if (int rc = check_handle(h)) return rc;
  std::lock_guard<std::mutex> lk(h->mu);
  h->profiling = enable != 0;
  for (int k = 0; k < 3; ++k) {
    h->prof_ms[k] = 0;
    h->prof_launches[k] = 0;
  }
  return FECL_OK;
}

int fecl_ldpc_profile_read(fecl_ldpc* h, double* ms, int64_t* launches) {
  if (int rc = check_handle(h)) return rc;
  if (!ms || !launches) return fail(FECL_INVALID_ARGUMENT, "null output");
  std::lock_guard<std::mutex> lk(h->mu);
  for (int k = 0; k < 3; ++k) {
    ms[k] = h->prof_ms[k];
    launches[k] = h->prof_launches[k];
  }
  return FECL_OK;
}

int fecl_ldpc_set_layout(fecl_ldpc* h, int32_t words) {
  if (int rc = check_handle(h)) return rc;
  if (words != 0 && words != 1 && words != 2 && words != 4) return fail(FECL_INVALID_ARGUMENT, "words must be 0, 1, 2 or 4");
  if (words) {
    Layout lay;
    lay.fmt = h->msg_format;
    lay.words = words;
    if (!layout_fits(lay, h->algorithm, h->code.max_row_degree, h->code.max_var_degree))
      return fail(FECL_NOT_SUPPORTED, "layout exceeds the register budget for this code's degrees");
  }
  std::lock_guard<std::mutex> lk(h->mu);
  h->words_pinned = words;
  choose_layout(h);
  return FECL_OK;
}

int fecl_ldpc_set_fused_syndrome(fecl_ldpc* h, int32_t enable) {
  if (int rc = check_handle(h)) return rc;
  std::lock_guard<std::mutex> lk(h->mu);
  h->fused_syndrome = enable != 0;
  return FECL_OK;
}

int fecl_ldpc_set_vars_per_wave(fecl_ldpc* h, int32_t vars) {
  if (int rc = check_handle(h)) return rc;
  if (vars < 0 || vars > 2) return fail(FECL_INVALID_ARGUMENT, "variables per wavefront must be 0, 1 or 2");
  if (vars == 2) {
    Layout lay = layout_of(h);
    lay.vpw = 2;
    if (!vpw_fits(lay, h->code.max_var_degree))
      return fail(FECL_NOT_SUPPORTED, "two variables per wavefront exceed the register budget for this layout");
  }
  std::lock_guard<std::mutex> lk(h->mu);
  h->vpw_pinned = vars;
  choose_layout(h);
  return FECL_OK;
}

int fecl_ldpc_get_vars_per_wave(const fecl_ldpc* h, int32_t* vars) {
  if (int rc = check_handle(h)) return rc;
  if (vars) *vars = h->vpw;
  return FECL_OK;
}

int fecl_ldpc_set_compaction(fecl_ldpc* h, int32_t enable) {
  if (int rc = check_handle(h)) return rc;
  std::lock_guard<std::mutex> lk(h->mu);
  h->compaction = enable != 0;
  return FECL_OK;
}

int fecl_ldpc_set_cache_policy(fecl_ldpc* h, int32_t policy) {
  if (int rc = check_handle(h)) return rc;
  if (policy < -1 || policy > 3) return fail(FECL_INVALID_ARGUMENT, "cache policy must be -1 (automatic) or 0..3");
  std::lock_guard<std::mutex> lk(h->mu);
  h->cache_policy = policy;
  return FECL_OK;
}

int fecl_ldpc_set_pass_frames(fecl_ldpc* h, int64_t frames) {
  if (int rc = check_handle(h)) return rc;
  if (frames < 0) return fail(FECL_INVALID_ARGUMENT, "frames per pass must be >= 0");
  std::lock_guard<std::mutex> lk(h->mu);
  h->pass_frames = frames;
  return FECL_OK;
}

int fecl_ldpc_get_layout(const fecl_ldpc* h, int32_t* words, int32_t* frames_per_tile) {
  if (int rc = check_handle(h)) return rc;
  const Layout lay = layout_of(h);
  if (words) *words = lay.words;
  if (frames_per_tile) *frames_per_tile = lay.tw();
  return FECL_OK;
}

int fecl_ldpc_autotune(fecl_ldpc* h, int64_t frames, int32_t probe_iterations, int32_t* best_words, double* ms) {
  if (int rc = check_handle(h)) return rc;
  if (frames < 1 || probe_iterations < 1) return fail(FECL_INVALID_ARGUMENT, "frames and probe_iterations must be >= 1");
  const Code& c = h->code;
  HIP_TRY(hipSetDevice(h->device));
  // a pick kept for this GPU, code, algorithm, message format and batch size is pinned as is
  tune::configure_from_env();
  hipDeviceProp_t prop{};
  HIP_TRY(hipGetDeviceProperties(&prop, h->device));
  const std::string key = tune::make_key(prop.gcnArchName, prop.multiProcessorCount, c, h->algorithm, h->msg_format,
                                         frames);
  {
    tune::Pick kept;
    if (tune::lookup(key, &kept)) {
      Layout lay;
      lay.fmt = h->msg_format;
      lay.words = kept.words;
      lay.vpw = kept.vpw;
      if (layout_fits(lay, h->algorithm, c.max_row_degree, c.max_var_degree) && vpw_fits(lay, c.max_var_degree)) {
        h->words_pinned = kept.words;
        h->vpw_pinned = kept.vpw;
        choose_layout(h);
        if (ms)
          for (int i = 0; i < 6; ++i) ms[i] = kept.ms[i];
        if (best_words) *best_words = h->words;
        return FECL_OK;
      }
    }
  }
  // Noise-only LLRs: no frame converges, so every probe iteration runs for every frame.
  std::vector<float> host((size_t)frames * c.N);
  uint64_t x = 0x9E3779B97F4A7C15ull;
  for (auto& v : host) {
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    v = (float)((int64_t)(x >> 40) - (1 << 23)) / (float)(1 << 22);
  }
  DevBuf llr, bits;
  HIP_TRY(llr.ensure(host.size() * 4));
  HIP_TRY(bits.ensure((size_t)frames * c.K));
  HIP_TRY(hipMemcpy(llr.p, host.data(), host.size() * 4, hipMemcpyHostToDevice));
  const int32_t saved_iters = h->iterations, saved_pin = h->words_pinned, saved_vpin = h->vpw_pinned;
  const bool saved_prof = h->profiling;
  h->iterations = probe_iterations;
  h->profiling = false;
  hipStream_t st = nullptr;
  HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  double best = 1e30;
  int32_t best_w = 0, best_v = 0;
  const int cand[3] = {1, 2, 4};
  int rc = FECL_OK;
  for (int i = 0; i < 6 && rc == FECL_OK; ++i) {
    Layout lay;
    lay.fmt = h->msg_format;
    lay.words = cand[i / 2];
    lay.vpw = 1 + i % 2;
    if (ms) ms[i] = -1.0;
    if (!layout_fits(lay, h->algorithm, c.max_row_degree, c.max_var_degree) || !vpw_fits(lay, c.max_var_degree))
      continue;
    h->words_pinned = lay.words;
    h->vpw_pinned = lay.vpw;
    choose_layout(h);
    // The first run sizes the workspace; the fastest of the next three counts (layouts within
    // a few per cent of each other would otherwise be ranked by run-to-run noise).
    float t = 0.0f;
    for (int rep = 0; rep < 4 && rc == FECL_OK; ++rep) {
      if (hipEventRecord(e0, st) != hipSuccess) {
        rc = fail(FECL_DEVICE_ERROR, "autotune: hipEventRecord failed");
        break;
      }
      rc = device_decode(h, llr.as<float>(), nullptr, nullptr, frames, bits.as<uint8_t>(), nullptr, nullptr,
                         nullptr, nullptr, nullptr, st);
      float r = 0.0f;
      if (rc == FECL_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                            hipEventElapsedTime(&r, e0, e1) != hipSuccess))
        rc = fail(FECL_DEVICE_ERROR, "autotune: event timing failed");
      if (rep == 1 || (rep > 1 && r < t)) t = r;
    }
    if (ms) ms[i] = t;
    if (rc == FECL_OK && t < best) { best = t; best_w = lay.words; best_v = lay.vpw; }
  }
  h->iterations = saved_iters;
  h->profiling = saved_prof;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  if (rc != FECL_OK || !best_w) {
    h->words_pinned = saved_pin;
    h->vpw_pinned = saved_vpin;
    choose_layout(h);
    return rc;
  }
  h->words_pinned = best_w;
  h->vpw_pinned = best_v;
  choose_layout(h);
  if (best_words) *best_words = h->words;
  tune::Pick pick;
  pick.words = best_w;
  pick.vpw = best_v;
  if (ms)
    for (int i = 0; i < 6; ++i) pick.ms[i] = ms[i];
  // a cache file that cannot be appended does not undo the pick (stats count the failure)
  std::string err;
  (void)tune::store(key, pick, &err);
  return FECL_OK;
}

int fecl_autotune_cache_config(int32_t mode, const char* path) {
  std::string err;
  if (!tune::configure(mode, path, &err)) return fail(FECL_INVALID_ARGUMENT, err);
  return FECL_OK;
}

int fecl_autotune_cache_stats(int64_t* hits, int64_t* misses, int64_t* entries, int64_t* skipped_lines,
                              int64_t* write_failures) {
  tune::stats(hits, misses, entries, skipped_lines, write_failures);
  return FECL_OK;
}

int fecl_autotune_cache_clear(void) {
  tune::clear();
  return FECL_OK;
}

}  // extern "C"

// A captured multi-code batch: one branch per code, all launched as one hipGraph.
struct fecl_ldpc_graph {
  std::vector<fecl_ldpc*> handles;
  std::vector<std::unique_ptr<Workspace>> ws;
  std::vector<std::unique_ptr<OptsCopy>> opts;   // each branch's scaling tables, as at creation
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  hipStream_t capture = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  PersistBatch* persist = nullptr;   // the persistent form's workspaces and control block
  ~fecl_ldpc_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : events) (void)hipEventDestroy(e);
    for (auto st : streams) (void)hipStreamDestroy(st);
    if (capture) (void)hipStreamDestroy(capture);
    if (persist) persist_destroy(persist);
  }
};

extern "C" {

int fecl_ldpc_graph_create(fecl_ldpc* const* handles, const int64_t* frames, int32_t count, const float* const* llr,
                           uint8_t* const* msg, int32_t* const* iterations, fecl_ldpc_graph** out) {
  if (!out || !handles || !frames || !llr || !msg || count < 1) return fail(FECL_INVALID_ARGUMENT, "invalid arguments");
  *out = nullptr;
  std::unique_ptr<fecl_ldpc_graph> g(new fecl_ldpc_graph());
  const int dev = handles[0]->device;
  HIP_TRY(hipSetDevice(dev));
  for (int32_t i = 0; i < count; ++i) {
    if (!handles[i] || frames[i] < 1 || !llr[i] || !msg[i]) return fail(FECL_INVALID_ARGUMENT, "invalid branch");
    if (handles[i]->device != dev) return fail(FECL_INVALID_ARGUMENT, "all codes must live on one device");
  }
  HIP_TRY(hipStreamCreateWithFlags(&g->capture, hipStreamNonBlocking));
  Layout common;
  if (persist_eligible(handles, count, &common)) {
    // one persistent kernel for every code (ldpc_persist.hip), preceded by the counter reset
    // and each code's ingest
    std::vector<DevOpts> opts;
    for (int32_t i = 0; i < count; ++i) {
      g->handles.push_back(handles[i]);
      std::unique_ptr<OptsCopy> oc(new OptsCopy());
      if (int rc = copy_opts(handles[i], oc.get())) return rc;
      opts.push_back(oc->opts);
      g->opts.push_back(std::move(oc));
    }
    if (int rc = persist_create(handles, frames, count, llr, msg, iterations, common, opts.data(), &g->persist))
      return rc;
    HIP_TRY(hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal));
    const int rc = persist_enqueue(g->persist, handles, g->capture);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(g->capture, &graph);
    g->graph = graph;
    if (rc != FECL_OK) return rc;
    HIP_TRY(e2);
    HIP_TRY(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g.release();
    return FECL_OK;
  }
  hipEvent_t fork = nullptr;
  HIP_TRY(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
  g->events.push_back(fork);
  // Size every workspace before capture: nothing may allocate while the graph is recorded.
  for (int32_t i = 0; i < count; ++i) {
    fecl_ldpc* h = handles[i];
    g->handles.push_back(h);
    std::unique_ptr<Workspace> ws(new Workspace());
    DevState s;
    DevOut o;
    if (int rc = prepare(h, ws.get(), layout_of(h), frames[i], false, &s, &o)) return rc;
    if (h->compaction)
      if (int rc = prepare_graph_compaction(h, ws.get(), frames[i])) return rc;
    g->ws.push_back(std::move(ws));
    std::unique_ptr<OptsCopy> oc(new OptsCopy());
    if (int rc = copy_opts(h, oc.get())) return rc;
    g->opts.push_back(std::move(oc));
    hipStream_t st = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    g->streams.push_back(st);
    hipEvent_t join = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    g->events.push_back(join);
  }
  HIP_TRY(hipStreamBeginCapture(g->capture, hipStreamCaptureModeThreadLocal));
  int rc = FECL_OK;
  hipError_t e = hipEventRecord(fork, g->capture);
  for (int32_t i = 0; i < count && rc == FECL_OK && e == hipSuccess; ++i) {
    e = hipStreamWaitEvent(g->streams[i], fork, 0);
    if (e != hipSuccess) break;
    BatchIO io{};
    io.llr = llr[i];
    io.res.msg_bits = msg[i];
    io.res.iterations = iterations ? iterations[i] : nullptr;
    rc = run_passes(handles[i], g->ws[i].get(), g->streams[i], io, frames[i], 0, /*capture=*/true, &g->opts[i]->opts);
    if (rc != FECL_OK) break;
    e = hipEventRecord(g->events[i + 1], g->streams[i]);
    if (e == hipSuccess) e = hipStreamWaitEvent(g->capture, g->events[i + 1], 0);
  }
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(g->capture, &graph);
  g->graph = graph;
  if (rc != FECL_OK) return rc;
  HIP_TRY(e);
  HIP_TRY(e2);
  HIP_TRY(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  *out = g.release();
  return FECL_OK;
}

int fecl_ldpc_graph_launch(fecl_ldpc_graph* g, void* stream) {
  if (!g) return fail(FECL_INVALID_ARGUMENT, "null graph");
  HIP_TRY(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return FECL_OK;
}

int fecl_ldpc_graph_check(fecl_ldpc_graph* g) {
  if (!g) return fail(FECL_INVALID_ARGUMENT, "null graph");
  if (!g->persist) return FECL_OK;
  int32_t error = 0;
  if (int rc = persist_error(g->persist, &error)) return rc;
  if (error) return fail(FECL_DEVICE_ERROR, "persistent decode stopped on its watchdog");
  return FECL_OK;
}

int fecl_ldpc_graph_is_persistent(const fecl_ldpc_graph* g, int32_t* persistent) {
  if (!g || !persistent) return fail(FECL_INVALID_ARGUMENT, "invalid arguments");
  *persistent = g->persist ? 1 : 0;
  return FECL_OK;
}

int fecl_ldpc_set_graph_persistent(fecl_ldpc* h, int32_t enable) {
  if (int rc = check_handle(h)) return rc;
  std::lock_guard<std::mutex> lk(h->mu);
  h->persistent = enable != 0;
  return FECL_OK;
}

int fecl_ldpc_graph_destroy(fecl_ldpc_graph* g) {
  delete g;
  return FECL_OK;
}

int fecl_ldpc_trim(fecl_ldpc* h) {
  if (int rc = check_handle(h)) return rc;
  std::lock_guard<std::mutex> lk(h->mu);
  h->pool.clear();
  return FECL_OK;
}

}  // extern "C"
