// Continuous-batching decode server (C ABI in include/fecl_ldpc.h, fecl_ldpc_server_*).
//
// FeCl has no serving layer: Codec::decode blocks on one batch (src/Codec.h:210-238).  Here
// each code owns a pool of device frame slots whose message state stays resident across
// iterations.  A worker thread runs rounds, each a few steps for every pool:
//   admit   queued frames into free slots (their LLRs are copied into the slots' tile rows;
//           LLRs already in device memory are read in place)
//   step    check-node, variable-node (+ outputs of finishing frames), syndrome/control,
//           repeated steps_per_round times with no host work in between
//   retire  frames that reached DONE: read their hard bits and iteration counts, free the slots
// with the device kept one round ahead of the host's bookkeeping.  Requests that arrive while
// others are on the GPU join at a round boundary, and a frame's arithmetic is exactly that of
// a batch decode (no cross-frame state).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <queue>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ldpc_admission.h"
#include "ldpc_engine.h"

namespace fecl {
namespace engine {

using Clock = std::chrono::steady_clock;

// Steps per round unless fecl_ldpc_server_set_round_steps says otherwise, and the limit.
constexpr int kDefaultRoundSteps = 1;
constexpr int kMaxRoundSteps = 64;

struct Request {
  uint64_t id = 0;
  int pool = 0;
  const float* llr = nullptr;   // [frames][N], host or device memory
  bool pinned = false;          // llr is page-locked: admission copies it straight to the device
  bool device = false;          // llr is device memory: admission kernels read it in place
  int64_t frames = 0;
  uint8_t* msg = nullptr;       // host [frames][K]
  int32_t* iters = nullptr;     // host [frames] or null
  int64_t admitted = 0, done = 0;
  Clock::time_point t_submit, t_done;
  bool finished = false;
};

struct PinnedBuf {
  void* p = nullptr;
  void* dev = nullptr;   // the same memory as the device addresses it (kernels read it in place)
  size_t bytes = 0;
  hipError_t ensure(size_t n) {
    if (n <= bytes) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = dev = nullptr;
    bytes = 0;
    hipError_t e = hipHostMalloc(&p, n, hipHostMallocMapped);
    if (e == hipSuccess) e = hipHostGetDevicePointer(&dev, p, 0);
    if (e == hipSuccess) bytes = n;
    return e;
  }
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

struct Pool {
  fecl_ldpc* h = nullptr;
  Layout lay;
  OptsCopy oc;              // the handle's options at creation, scaling tables included
  int32_t iterations = 1;
  DevOpts opts{};
  Workspace ws;
  DevState s{};
  DevOut out{};
  int64_t nslots = 0;
  // free slots of each tile and the tile's occupied slots (see choose_admissions)
  std::vector<std::vector<int32_t>> tile_free;
  std::vector<int32_t> tile_used;
  std::vector<std::pair<Request*, int64_t>> owner;  // slot -> (request, frame)
  int64_t live = 0;
  std::deque<Request*> queue;
  std::vector<int64_t> admit_round;                 // slot -> round its current frame was admitted in
  DevBuf d_stage[2], d_admit[2], d_ext_bits[2], d_ext_iters[2];
  // host buffers, two of each: the worker fills or reads one while the device uses the other
  PinnedBuf h_stage[2], h_admit[2], h_ext_slots[2], h_ext_bits[2], h_ext_iters[2], h_phase[2];
  std::vector<std::pair<Request*, int64_t>> ext_owner[2];   // retired entries awaiting their results
  ~Pool() {
    for (DevBuf* b : {&d_stage[0], &d_stage[1], &d_admit[0], &d_admit[1], &d_ext_bits[0], &d_ext_bits[1],
                      &d_ext_iters[0], &d_ext_iters[1]})
      b->release();
  }
};

}  // namespace engine
}  // namespace fecl

using namespace fecl;
using namespace fecl::engine;

struct fecl_ldpc_server {
  std::vector<std::unique_ptr<Pool>> pools;
  std::mutex mu;                       // queues, requests, statistics
  std::condition_variable work_cv, done_cv;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
  uint64_t next_id = 1;
  bool stop = false;
  int error = FECL_OK;
  std::string error_msg;
  int64_t frames_done = 0, steps = 0, queued_frames = 0;
  int steps_per_round = kDefaultRoundSteps;   // read by the worker at each round (under mu)
  hipStream_t st = nullptr;      // admissions' kernels, steps, retirements' extracts
  hipStream_t copy = nullptr;    // admissions' host-to-device copies, overlapping the steps
  hipStream_t out = nullptr;     // results to the host, overlapping the steps
  int device = 0;
  std::thread worker;
  ~fecl_ldpc_server() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    work_cv.notify_all();
    if (worker.joinable()) worker.join();
    (void)hipSetDevice(device);
    for (hipStream_t x : {st, copy, out})
      if (x) (void)hipStreamSynchronize(x);
    pools.clear();
    for (hipStream_t x : {st, copy, out})
      if (x) (void)hipStreamDestroy(x);
  }
};

namespace {

// Frames each pool may admit per step of a round: admissions of a round are capped at this
// times the round's steps (and the pool's free slots).
constexpr int kAdmitPerStep = 128;

// The worker keeps the device one round ahead of the host.  Round e = admissions chosen from
// the free slots, one step of every pool holding frames, and a copy of every slot's phase.
// While the device runs round e, the host reads the phases of round e - 1 and retires the
// frames that were DONE by then: an extract of their bits and iteration counts is queued
// behind round e, and their slots may take new frames from round e + 1 on (stream order runs
// the extract first).  The results of that extract are handed out in the next iteration.
// Host buffers come in pairs indexed by the round's parity, so the host never touches a
// buffer the device may still be using.
struct Events {
  hipEvent_t round[2] = {nullptr, nullptr};     // round e's phases in host memory (compute stream)
  hipEvent_t ext[2] = {nullptr, nullptr};       // results of the extract queued after round q (out stream)
  hipEvent_t copied[2] = {nullptr, nullptr};    // admissions' copies of round e (copy stream)
  hipEvent_t admitted[2] = {nullptr, nullptr};  // admission kernels of round e (compute stream)
  hipEvent_t extracted[2] = {nullptr, nullptr}; // extracts queued after round q (compute stream)
  hipError_t create() {
    for (hipEvent_t* e : all())
      if (hipError_t r = hipEventCreateWithFlags(e, hipEventDisableTiming)) return r;
    return hipSuccess;
  }
  std::vector<hipEvent_t*> all() {
    std::vector<hipEvent_t*> v;
    for (int i = 0; i < 2; ++i)
      for (hipEvent_t* e : {&round[i], &ext[i], &copied[i], &admitted[i], &extracted[i]}) v.push_back(e);
    return v;
  }
  ~Events() {
    for (hipEvent_t* e : all())
      if (*e) (void)hipEventDestroy(*e);
  }
};

struct Admit { Request* req; int64_t f0, n; };

// Host side of the admissions of round e: queued frames into free slots of the tiles that
// admission_order (ldpc_admission.h) lists, lowest slots first; free slots of other tiles wait
// for a later round.  Returns whether any pool holds frames (and so steps).
constexpr int kFillTiles = 2;

bool choose_admissions(fecl_ldpc_server* sv, int64_t e, int steps, std::vector<std::vector<Admit>>* admits) {
  bool any = false;
  std::vector<int> order;
  std::vector<int64_t> free_count, used_count;
  std::lock_guard<std::mutex> lk(sv->mu);
  for (size_t pi = 0; pi < sv->pools.size(); ++pi) {
    Pool& P = *sv->pools[pi];
    (*admits)[pi].clear();
    const int64_t cap = std::min<int64_t>(P.nslots, (int64_t)kAdmitPerStep * steps);
    int64_t n = 0;
    int32_t* slots = P.h_admit[e & 1].as<int32_t>();
    if (!P.queue.empty()) {
      const int T = (int)P.tile_free.size();
      int64_t want = 0;   // frames the round admits, if the tiles have room
      for (const Request* q : P.queue) {
        want += q->frames - q->admitted;
        if (want >= cap) break;
      }
      free_count.resize(T);
      used_count.resize(T);
      for (int t = 0; t < T; ++t) {
        free_count[t] = (int64_t)P.tile_free[t].size();
        used_count[t] = P.tile_used[t];
      }
      admission_order(free_count, used_count, std::min(want, cap), kFillTiles, &order);
      for (int t : order) {
        std::vector<int32_t>& fl = P.tile_free[t];
        std::sort(fl.begin(), fl.end(), std::greater<int32_t>());   // lowest at the back
        while (!fl.empty() && !P.queue.empty() && n < cap) {
          Request* q = P.queue.front();
          const int64_t k = std::min<int64_t>({(int64_t)fl.size(), q->frames - q->admitted, cap - n});
          for (int64_t i = 0; i < k; ++i) {
            const int32_t slot = fl.back();
            fl.pop_back();
            slots[n + i] = slot;
            P.owner[slot] = {q, q->admitted + i};
            P.admit_round[slot] = e;
          }
          P.tile_used[t] += (int32_t)k;
          (*admits)[pi].push_back({q, q->admitted, k});
          q->admitted += k;
          sv->queued_frames -= k;
          n += k;
          if (q->admitted == q->frames) P.queue.pop_front();
        }
        if (P.queue.empty() || n >= cap) break;
      }
    }
    P.live += n;
    any |= P.live > 0;
  }
  return any;
}

// Device side of round e: admission copies of host frames (on the copy stream, so they
// overlap the steps of round e - 1 still running) and admission kernels, `steps` steps per
// pool, the phase snapshot.  Buffers of parity e & 1 were last used by round e - 2, which the host
// has already seen complete, so they may be refilled or grown.
int enqueue_round(fecl_ldpc_server* sv, Events* ev, int64_t e, int steps, const std::vector<std::vector<Admit>>& admits) {
  const int b = (int)(e & 1);
  bool any_admit = false, any_copy = false;
  for (const auto& a : admits) any_admit |= !a.empty();
  if (any_admit) {
    if (e >= 2) HIP_TRY(hipStreamWaitEvent(sv->copy, ev->admitted[b], 0));
    for (size_t pi = 0; pi < sv->pools.size(); ++pi) {
      Pool& P = *sv->pools[pi];
      const int64_t N = P.h->code.N;
      int64_t n = 0, host = 0, pageable = 0;
      for (const Admit& a : admits[pi]) {
        n += a.n;
        if (!a.req->device) host += a.n;
        if (!a.req->device && !a.req->pinned) pageable += a.n;
      }
      if (!n) continue;
      // host frames are staged at their position among the round's admissions
      if (host) {
        if (P.d_stage[b].bytes < (size_t)(n * N * 4)) {
          HIP_TRY(hipDeviceSynchronize());   // growing: nothing may still read the old buffer
          HIP_TRY(P.d_stage[b].ensure((size_t)(n * N * 4)));
        }
        if (pageable) HIP_TRY(P.h_stage[b].ensure((size_t)(pageable * N * 4)));
      }
      int64_t off = 0, src = 0;
      for (const Admit& a : admits[pi]) {
        if (a.req->pinned) {   // page-locked request memory: copied straight to the device
          HIP_TRY(hipMemcpyAsync(P.d_stage[b].as<float>() + off * N, a.req->llr + a.f0 * N, (size_t)(a.n * N) * 4,
                                 hipMemcpyHostToDevice, sv->copy));
          any_copy = true;
        } else if (!a.req->device) {   // pageable: packed into pinned staging, then copied
          std::memcpy(P.h_stage[b].as<float>() + src * N, a.req->llr + a.f0 * N, (size_t)(a.n * N) * 4);
          HIP_TRY(hipMemcpyAsync(P.d_stage[b].as<float>() + off * N, P.h_stage[b].as<float>() + src * N,
                                 (size_t)(a.n * N) * 4, hipMemcpyHostToDevice, sv->copy));
          src += a.n;
          any_copy = true;
        }
        off += a.n;
      }
      HIP_TRY(hipMemcpyAsync(P.d_admit[b].p, P.h_admit[b].p, (size_t)n * 4, hipMemcpyHostToDevice, sv->copy));
      any_copy = true;
    }
    if (any_copy) {
      HIP_TRY(hipEventRecord(ev->copied[b], sv->copy));
      HIP_TRY(hipStreamWaitEvent(sv->st, ev->copied[b], 0));
    }
    for (size_t pi = 0; pi < sv->pools.size(); ++pi) {
      Pool& P = *sv->pools[pi];
      const uint8_t init = (uint8_t)((P.iterations == 1 ? PH_FINAL : PH_RUN) | PH_FRESH);
      // one launch per run of staged host frames, one per device-resident chunk
      int64_t off = 0, run0 = -1;
      auto flush = [&](int64_t end) -> int {
        if (run0 >= 0 && end > run0)
          HIP_TRY(kernels_admit(P.lay, P.h->dc, P.s, P.d_stage[b].as<float>() + run0 * P.h->code.N,
                                P.d_admit[b].as<int32_t>() + run0, (int)(end - run0), init, sv->st));
        run0 = -1;
        return FECL_OK;
      };
      for (const Admit& a : admits[pi]) {
        if (a.req->device) {
          if (int rc = flush(off)) return rc;
          HIP_TRY(kernels_admit(P.lay, P.h->dc, P.s, a.req->llr + a.f0 * P.h->code.N, P.d_admit[b].as<int32_t>() + off,
                                (int)a.n, init, sv->st));
        } else if (run0 < 0) {
          run0 = off;
        }
        off += a.n;
      }
      if (int rc = flush(off)) return rc;
    }
    HIP_TRY(hipEventRecord(ev->admitted[b], sv->st));
  }
  for (int j = 0; j < steps; ++j) {
    for (auto& pp : sv->pools) {
      Pool& P = *pp;
      if (!P.live) continue;
      HIP_TRY(kernels_check(P.lay, P.h->dc, P.s, P.opts, P.h->code.max_row_degree, sv->st));
      HIP_TRY(kernels_variable(P.lay, P.h->dc, P.s, P.out, P.h->code.max_var_degree, P.opts.fused_syn,
                               (P.opts.nt_rows & 2) != 0, sv->st));
      HIP_TRY(kernels_syndrome_control(P.lay, P.h->dc, P.s, P.iterations, P.opts.fused_syn, -1,
                                       P.h->code.max_row_degree, sv->st));
    }
  }
  // every pool's phases as this round leaves them, written by one kernel in stream order into
  // mapped page-locked memory: the next round's admissions and steps, queued behind it, cannot
  // change them while they are copied (a copy on another stream would race with them)
  ByteCopies snap{};
  for (auto& pp : sv->pools) {
    Pool& P = *pp;
    snap.item[snap.count++] = ByteCopy{P.s.phase, static_cast<uint8_t*>(P.h_phase[b].dev), P.nslots};
  }
  HIP_TRY(kernels_copy_bytes(snap, sv->st));
  HIP_TRY(hipEventRecord(ev->round[b], sv->st));
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->steps += steps;
  return FECL_OK;
}

// Round q has completed: retire the frames its phases show DONE and queue their extract.
int retire_round(fecl_ldpc_server* sv, Events* ev, int64_t q, bool* queued) {
  const int b = (int)(q & 1);
  bool any = false;
  *queued = false;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    for (auto& pp : sv->pools) {
      Pool& P = *pp;
      P.ext_owner[b].clear();
      const uint8_t* ph = P.h_phase[b].as<uint8_t>();
      int32_t* list = P.h_ext_slots[b].as<int32_t>();
      for (int64_t slot = 0; slot < P.nslots; ++slot) {
        // a frame admitted after round q is not the one these phases describe
        if (!P.owner[slot].first || P.admit_round[slot] > q || (ph[slot] & PH_MASK) != PH_DONE) continue;
        list[P.ext_owner[b].size()] = (int32_t)slot;
        P.ext_owner[b].push_back(P.owner[slot]);
        P.owner[slot] = {nullptr, 0};
        const int t = (int)(slot / P.lay.tw());
        P.tile_free[t].push_back((int32_t)slot);
        --P.tile_used[t];
        --P.live;
      }
      any |= !P.ext_owner[b].empty();
    }
  }
  if (!any) return FECL_OK;
  // the extract kernels read the slot lists in place from page-locked memory; their results
  // (buffers of parity b, free since the host handed out round q - 2's) go back on the out
  // stream while the next round steps
  for (auto& pp : sv->pools) {
    Pool& P = *pp;
    const int n = (int)P.ext_owner[b].size();
    if (!n) continue;
    HIP_TRY(kernels_extract(P.lay, P.h->dc, P.s, P.out, static_cast<const int32_t*>(P.h_ext_slots[b].dev), n,
                            P.d_ext_bits[b].as<uint8_t>(), P.d_ext_iters[b].as<int32_t>(), sv->st));
  }
  HIP_TRY(hipEventRecord(ev->extracted[b], sv->st));
  HIP_TRY(hipStreamWaitEvent(sv->out, ev->extracted[b], 0));
  for (auto& pp : sv->pools) {
    Pool& P = *pp;
    const int n = (int)P.ext_owner[b].size();
    if (!n) continue;
    const int64_t K = P.h->code.K;
    HIP_TRY(hipMemcpyAsync(P.h_ext_bits[b].p, P.d_ext_bits[b].p, (size_t)n * K, hipMemcpyDeviceToHost, sv->out));
    HIP_TRY(hipMemcpyAsync(P.h_ext_iters[b].p, P.d_ext_iters[b].p, (size_t)n * 4, hipMemcpyDeviceToHost, sv->out));
  }
  HIP_TRY(hipEventRecord(ev->ext[b], sv->out));
  *queued = true;
  return FECL_OK;
}

// The extract queued for round q has finished: copy its results into the requests.
void hand_out(fecl_ldpc_server* sv, int64_t q) {
  const int b = (int)(q & 1);
  const Clock::time_point now = Clock::now();
  bool notify = false;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    for (auto& pp : sv->pools) {
      Pool& P = *pp;
      const int64_t K = P.h->code.K;
      for (size_t i = 0; i < P.ext_owner[b].size(); ++i) {
        Request* r = P.ext_owner[b][i].first;
        const int64_t f = P.ext_owner[b][i].second;
        std::memcpy(r->msg + f * K, P.h_ext_bits[b].as<uint8_t>() + (int64_t)i * K, (size_t)K);
        if (r->iters) r->iters[f] = P.h_ext_iters[b].as<int32_t>()[i];
        ++sv->frames_done;
        if (++r->done == r->frames) {
          r->finished = true;
          r->t_done = now;
          notify = true;
        }
      }
      P.ext_owner[b].clear();
    }
  }
  if (notify) sv->done_cv.notify_all();
}

int server_run(fecl_ldpc_server* sv) {
  Events ev;
  HIP_TRY(ev.create());
  std::vector<std::vector<Admit>> admits(sv->pools.size());
  int64_t enqueued = 0, processed = 0;   // rounds enqueued / rounds whose phases were read
  int64_t ext_round = -1;                // round whose extract results are still to be handed out
  for (;;) {
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      if (sv->stop) return FECL_OK;
    }
    int steps = 1;
    {
      std::lock_guard<std::mutex> lk(sv->mu);
      steps = sv->steps_per_round;
    }
    const bool live = choose_admissions(sv, enqueued, steps, &admits);
    if (live) {
      if (int rc = enqueue_round(sv, &ev, enqueued, steps, admits)) return rc;
      ++enqueued;
    }
    // read every completed round but the newest (which keeps the device busy meanwhile)
    while (processed < enqueued - (live ? 1 : 0)) {
      HIP_TRY(hipEventSynchronize(ev.round[processed & 1]));
      if (ext_round >= 0) {
        HIP_TRY(hipEventSynchronize(ev.ext[ext_round & 1]));
        hand_out(sv, ext_round);
        ext_round = -1;
      }
      bool queued = false;
      if (int rc = retire_round(sv, &ev, processed, &queued)) return rc;
      if (queued) ext_round = processed;
      ++processed;
    }
    if (live) continue;
    if (ext_round >= 0) {   // nothing to step: finish the last extract, its slots are free now
      HIP_TRY(hipEventSynchronize(ev.ext[ext_round & 1]));
      hand_out(sv, ext_round);
      ext_round = -1;
      continue;
    }
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->work_cv.wait_for(lk, std::chrono::milliseconds(5), [&] {
      if (sv->stop) return true;
      for (auto& p : sv->pools)
        if (!p->queue.empty()) return true;
      return false;
    });
  }
}

void server_loop(fecl_ldpc_server* sv) {
  (void)hipSetDevice(sv->device);
  const int rc = server_run(sv);
  if (rc != FECL_OK) {
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->error = rc;
    sv->error_msg = fecl_last_error();
    sv->stop = true;
  }
  sv->done_cv.notify_all();
}

// Device bytes one slot needs in a pool of this layout (messages, channel, bookkeeping).
int64_t slot_bytes(const fecl_ldpc* h, const Layout& lay) {
  const Code& c = h->code;
  return c.E * 4 / lay.pw() + c.N * 4 + (c.N + c.K) * 8 / 64 + 8;
}

}  // namespace

extern "C" {

int fecl_ldpc_server_create(fecl_ldpc* const* handles, int32_t count, const int64_t* slots, int64_t memory_bytes,
                            fecl_ldpc_server** out) {
  if (!out || !handles || count < 1) return fail(FECL_INVALID_ARGUMENT, "invalid arguments");
  if (count > kMaxByteCopies) return fail(FECL_INVALID_ARGUMENT, "at most 32 codes per server");
  *out = nullptr;
  std::unique_ptr<fecl_ldpc_server> sv(new fecl_ldpc_server());
  sv->device = handles[0]->device;
  HIP_TRY(hipSetDevice(sv->device));
  // Admission by device memory: pools without an explicit slot count share the budget (by
  // default half of the free memory) in proportion to one slot's footprint.
  if (memory_bytes <= 0) {
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    memory_bytes = (int64_t)(free_b / 2);
  }
  int unset = 0;
  for (int32_t i = 0; i < count; ++i) {
    if (!handles[i] || handles[i]->device != sv->device) return fail(FECL_INVALID_ARGUMENT, "invalid handle");
    if (!slots || slots[i] <= 0) ++unset;
  }
  for (int32_t i = 0; i < count; ++i) {
    fecl_ldpc* h = handles[i];
    std::unique_ptr<Pool> P(new Pool());
    P->h = h;
    P->lay = layout_of(h);
    if (int rc = copy_opts(h, &P->oc)) return rc;
    P->opts = P->oc.opts;
    P->iterations = h->iterations;
    const int64_t tw = P->lay.tw();
    int64_t want = (slots && slots[i] > 0) ? slots[i] : (memory_bytes / unset) / slot_bytes(h, P->lay);
    int64_t tiles = std::max<int64_t>(1, want / tw);
    P->nslots = tiles * tw;
    P->opts.nt_rows = nt_rows_for(h, P->lay, tiles);
    if (int rc = prepare(h, &P->ws, P->lay, P->nslots, false, &P->s, &P->out)) return rc;
    HIP_TRY(kernels_fill_i32(P->s.iter, P->nslots, 0, nullptr));
    HIP_TRY(hipMemset(P->s.phase, PH_FREE, (size_t)P->nslots));
    HIP_TRY(hipMemset(P->s.tile_flags, 0, (size_t)tiles * 4));
    HIP_TRY(hipMemset(P->s.fail, 0, (size_t)tiles * P->lay.fpl() * 8));
    const Code& c = h->code;
    // staging for host frames grows with the rounds' admissions; slot lists fit every slot
    for (int b = 0; b < 2; ++b) {
      HIP_TRY(P->d_stage[b].ensure((size_t)std::min<int64_t>(P->nslots, kAdmitPerStep) * c.N * 4));
      HIP_TRY(P->d_admit[b].ensure((size_t)P->nslots * 4));
    }
    for (int b = 0; b < 2; ++b) {
      HIP_TRY(P->d_ext_bits[b].ensure((size_t)P->nslots * c.K));
      HIP_TRY(P->d_ext_iters[b].ensure((size_t)P->nslots * 4));
    }
    for (int b = 0; b < 2; ++b) {
      HIP_TRY(P->h_admit[b].ensure((size_t)P->nslots * 4));
      HIP_TRY(P->h_ext_slots[b].ensure((size_t)P->nslots * 4));
      HIP_TRY(P->h_ext_bits[b].ensure((size_t)P->nslots * c.K));
      HIP_TRY(P->h_ext_iters[b].ensure((size_t)P->nslots * 4));
      HIP_TRY(P->h_phase[b].ensure((size_t)P->nslots));
    }
    P->owner.assign((size_t)P->nslots, {nullptr, 0});
    P->admit_round.assign((size_t)P->nslots, 0);
    P->tile_free.assign((size_t)tiles, {});
    P->tile_used.assign((size_t)tiles, 0);
    for (int64_t s = 0; s < P->nslots; ++s) P->tile_free[(size_t)(s / tw)].push_back((int32_t)s);
    sv->pools.push_back(std::move(P));
  }
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipStreamCreateWithFlags(&sv->st, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&sv->copy, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&sv->out, hipStreamNonBlocking));
  fecl_ldpc_server* raw = sv.get();
  sv->worker = std::thread(server_loop, raw);
  *out = sv.release();
  return FECL_OK;
}

int fecl_ldpc_server_slots(const fecl_ldpc_server* sv, int32_t code, int64_t* slots) {
  if (!sv || !slots || code < 0 || code >= (int32_t)sv->pools.size()) return fail(FECL_INVALID_ARGUMENT, "invalid arguments");
  *slots = sv->pools[code]->nslots;
  return FECL_OK;
}

int fecl_ldpc_server_submit(fecl_ldpc_server* sv, int32_t code, const float* llr, int64_t frames, uint8_t* msg,
                            int32_t* iterations, uint64_t* ticket) {
  if (!sv || !ticket || code < 0 || code >= (int32_t)sv->pools.size()) return fail(FECL_INVALID_ARGUMENT, "invalid arguments");
  if (frames < 1 || !llr || !msg) return fail(FECL_INVALID_ARGUMENT, "Invalid size for parity");
  std::unique_ptr<Request> r(new Request());
  r->pool = code;
  r->llr = llr;
  r->frames = frames;
  r->msg = msg;
  r->iters = iterations;
  {
    hipPointerAttribute_t attr{};
    const bool known = hipPointerGetAttributes(&attr, llr) == hipSuccess;
    (void)hipGetLastError();   // a pageable pointer is not an error here
    r->pinned = known && attr.type == hipMemoryTypeHost;
    r->device = known && attr.type == hipMemoryTypeDevice;
    if (r->device && attr.device != sv->device)
      return fail(FECL_INVALID_ARGUMENT, "LLRs in device memory of another device");
  }
  r->t_submit = Clock::now();
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (sv->stop) return fail(sv->error ? sv->error : FECL_DEVICE_ERROR, sv->error_msg.empty() ? "server stopped" : sv->error_msg);
    r->id = sv->next_id++;
    *ticket = r->id;
    sv->pools[code]->queue.push_back(r.get());
    sv->queued_frames += frames;
    sv->requests[r->id] = std::move(r);
  }
  sv->work_cv.notify_one();
  return FECL_OK;
}

int fecl_ldpc_server_wait(fecl_ldpc_server* sv, uint64_t ticket, double timeout_s, int32_t* done, double* latency_s) {
  if (!sv || !done) return fail(FECL_INVALID_ARGUMENT, "invalid arguments");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto it = sv->requests.find(ticket);
  if (it == sv->requests.end()) return fail(FECL_INVALID_ARGUMENT, "unknown ticket");
  Request* r = it->second.get();
  auto pred = [&] { return r->finished || sv->stop; };
  if (timeout_s < 0) sv->done_cv.wait(lk, pred);
  else sv->done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s), pred);
  *done = r->finished ? 1 : 0;
  if (r->finished) {
    if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    sv->requests.erase(it);
    return FECL_OK;
  }
  if (sv->error) return fail(sv->error, sv->error_msg);
  return FECL_OK;
}

int fecl_ldpc_server_set_round_steps(fecl_ldpc_server* sv, int32_t steps) {
  if (!sv || steps < 1 || steps > kMaxRoundSteps) return fail(FECL_INVALID_ARGUMENT, "invalid arguments");
  std::lock_guard<std::mutex> lk(sv->mu);
  sv->steps_per_round = steps;
  return FECL_OK;
}

int fecl_ldpc_server_stats(fecl_ldpc_server* sv, int64_t* frames_done, int64_t* steps, int64_t* queued_frames) {
  if (!sv) return fail(FECL_INVALID_ARGUMENT, "null server");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (frames_done) *frames_done = sv->frames_done;
  if (steps) *steps = sv->steps;
  if (queued_frames) *queued_frames = sv->queued_frames;
  return FECL_OK;
}

int fecl_ldpc_server_destroy(fecl_ldpc_server* sv) {
  delete sv;
  return FECL_OK;
}

}  // extern "C"
