// The serving loop (fecl_amd/csrc/ldpc_server.cpp) on the CPU fake backend (fakedev/): its
// admission, stepping, retirement and hand-out logic, with no GPU, built plain, with
// AddressSanitizer + UndefinedBehaviorSanitizer and with ThreadSanitizer.
//   1. requests of two codes from several submitting threads, in pageable, page-locked and
//      device memory, at 1 and several steps per round: every result equals the fake decode
//   2. admission by device memory: pools sized from memory_bytes, within the device budget
//   3. device out of memory while the pools are created: FECL_OUT_OF_MEMORY, nothing leaked
//   4. device out of memory while running (staging that must grow): the error reaches wait and
//      submit, and the server still shuts down
//   5. destroy with requests still queued
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "fake_engine.h"
#include "ldpc_engine.h"

using namespace fecl;

static int g_failures = 0;
#define EXPECT(cond, ...)                                 \
  do {                                                    \
    if (!(cond)) {                                        \
      ++g_failures;                                       \
      std::printf("FAIL %s:%d: %s | ", __FILE__, __LINE__, #cond); \
      std::printf(__VA_ARGS__);                           \
      std::printf("\n");                                  \
    }                                                     \
  } while (0)

// A handle with only what the serving loop reads: the code's sizes and degrees and the options.
static std::unique_ptr<fecl_ldpc> make_handle(int64_t N, int64_t K, int64_t E, int32_t iterations, int32_t fmt) {
  std::unique_ptr<fecl_ldpc> h(new fecl_ldpc());
  h->code.N = N;
  h->code.K = K;
  h->code.M = N - K;
  h->code.E = E;
  h->code.max_row_degree = 8;
  h->code.max_var_degree = 4;
  h->dc.N = (int32_t)N;
  h->dc.K = (int32_t)K;
  h->dc.M = (int32_t)(N - K);
  h->dc.E = (int32_t)E;
  h->iterations = iterations;
  h->msg_format = fmt;
  h->algorithm = FECL_APPROXIMATE;
  return h;
}

enum Mem { PAGEABLE = 0, PINNED = 1, DEVICE = 2 };

struct Req {
  int code = 0;
  int64_t frames = 0;
  Mem mem = PAGEABLE;
  std::vector<float> llr;         // host copy (the expected results come from it)
  float* buf = nullptr;           // what was submitted
  std::vector<uint8_t> msg;
  std::vector<int32_t> iters;
  uint64_t ticket = 0;
};

static void fill_request(Req* r, const fecl_ldpc* h, std::mt19937_64& rng) {
  std::normal_distribution<float> nd(0.5f, 2.0f);
  const int64_t N = h->code.N;
  r->llr.resize((size_t)(r->frames * N));
  for (float& x : r->llr) x = nd(rng);
  r->msg.assign((size_t)(r->frames * h->code.K), 0xee);
  r->iters.assign((size_t)r->frames, -7);
  const size_t bytes = r->llr.size() * 4;
  if (r->mem == PAGEABLE) {
    r->buf = r->llr.data();
  } else if (r->mem == PINNED) {
    void* p = nullptr;
    hipHostMalloc(&p, bytes, 0);
    std::memcpy(p, r->llr.data(), bytes);
    r->buf = static_cast<float*>(p);
  } else {
    void* p = nullptr;
    hipMalloc(&p, bytes);
    hipMemcpy(p, r->llr.data(), bytes, hipMemcpyHostToDevice);
    r->buf = static_cast<float*>(p);
  }
}

static void free_request(Req* r) {
  if (r->mem == PINNED) hipHostFree(r->buf);
  if (r->mem == DEVICE) hipFree(r->buf);
  r->buf = nullptr;
}

static bool check_request(const Req& r, const fecl_ldpc* h) {
  const int64_t N = h->code.N, K = h->code.K;
  for (int64_t f = 0; f < r.frames; ++f) {
    const float* l = &r.llr[(size_t)(f * N)];
    if (r.iters[(size_t)f] != fake_iterations(l, h->iterations)) return false;
    for (int64_t v = 0; v < K; ++v)
      if (r.msg[(size_t)(f * K + v)] != (l[v] < 0.0f ? 1 : 0)) return false;
  }
  return true;
}

static void test_mixed_requests(int round_steps) {
  auto a = make_handle(96, 48, 288, 12, FECL_MSG_F32);
  auto b = make_handle(64, 40, 200, 5, FECL_MSG_F16);
  fecl_ldpc* hs[2] = {a.get(), b.get()};
  const int64_t slots[2] = {192, 256};
  fecl_ldpc_server* sv = nullptr;
  int rc = fecl_ldpc_server_create(hs, 2, slots, 0, &sv);
  EXPECT(rc == FECL_OK, "create: %s", fecl_last_error());
  if (rc) return;
  EXPECT(fecl_ldpc_server_set_round_steps(sv, round_steps) == FECL_OK, "round steps");
  int64_t n0 = 0, n1 = 0;
  fecl_ldpc_server_slots(sv, 0, &n0);
  fecl_ldpc_server_slots(sv, 1, &n1);
  EXPECT(n0 == 192 && n1 == 256, "slots %lld %lld", (long long)n0, (long long)n1);
  const int kThreads = 4, kPerThread = 30;
  std::vector<std::vector<Req>> reqs(kThreads);
  std::atomic<int> bad{0};
  std::vector<std::thread> th;
  for (int t = 0; t < kThreads; ++t) {
    th.emplace_back([&, t] {
      std::mt19937_64 rng(1000 + 17 * t + round_steps);
      reqs[t].resize(kPerThread);
      for (int i = 0; i < kPerThread; ++i) {
        Req& r = reqs[t][i];
        r.code = (int)(rng() % 2);
        r.frames = 1 + (int64_t)(rng() % 70);   // some requests exceed a round's admissions
        r.mem = (Mem)(rng() % 3);
        fill_request(&r, hs[r.code], rng);
        if (fecl_ldpc_server_submit(sv, r.code, r.buf, r.frames, r.msg.data(), r.iters.data(), &r.ticket) != FECL_OK)
          ++bad;
      }
      for (int i = 0; i < kPerThread; ++i) {
        Req& r = reqs[t][i];
        int32_t done = 0;
        double lat = -1;
        if (fecl_ldpc_server_wait(sv, r.ticket, 60.0, &done, &lat) != FECL_OK || !done || lat < 0) ++bad;
      }
    });
  }
  for (auto& x : th) x.join();
  EXPECT(bad == 0, "%d submits or waits failed", bad.load());
  int64_t total = 0, wrong = 0;
  for (auto& v : reqs)
    for (Req& r : v) {
      total += r.frames;
      if (!check_request(r, hs[r.code])) ++wrong;
      free_request(&r);
    }
  EXPECT(wrong == 0, "%lld requests differ from the fake decode (round steps %d)", (long long)wrong, round_steps);
  int64_t done_frames = 0, steps = 0, queued = 0;
  fecl_ldpc_server_stats(sv, &done_frames, &steps, &queued);
  EXPECT(done_frames == total && queued == 0 && steps > 0, "stats %lld/%lld queued %lld", (long long)done_frames,
         (long long)total, (long long)queued);
  std::printf("mixed requests, %d steps per round: %d requests, %lld frames, %lld steps\n", round_steps,
              kThreads * kPerThread, (long long)total, (long long)steps);
  fecl_ldpc_server_destroy(sv);
}

static void test_admission_by_memory() {
  auto a = make_handle(96, 48, 288, 6, FECL_MSG_F32);
  auto b = make_handle(64, 40, 200, 6, FECL_MSG_F16);
  fecl_ldpc* hs[2] = {a.get(), b.get()};
  const size_t before = fake_device_bytes_in_use();
  fecl_ldpc_server* sv = nullptr;
  const int64_t budget = 4 << 20;
  int rc = fecl_ldpc_server_create(hs, 2, nullptr, budget, &sv);
  EXPECT(rc == FECL_OK, "create: %s", fecl_last_error());
  if (rc) return;
  int64_t n0 = 0, n1 = 0;
  fecl_ldpc_server_slots(sv, 0, &n0);
  fecl_ldpc_server_slots(sv, 1, &n1);
  // each pool gets half of the budget in whole tiles (64 FP32 frames, 128 FP16 frames)
  EXPECT(n0 > 0 && n1 > 0 && n0 % 64 == 0 && n1 % 128 == 0, "slots %lld %lld", (long long)n0, (long long)n1);
  const size_t used = fake_device_bytes_in_use() - before;
  EXPECT(used <= (size_t)budget * 2, "device bytes %zu for a budget of %lld", used, (long long)budget);
  // memory_bytes = 0: half of the free device memory
  fecl_ldpc_server* sv2 = nullptr;
  fake_set_device_budget(fake_device_bytes_in_use() + (64u << 20));
  rc = fecl_ldpc_server_create(hs, 2, nullptr, 0, &sv2);
  EXPECT(rc == FECL_OK, "create with free memory: %s", fecl_last_error());
  int64_t m0 = 0;
  if (sv2) fecl_ldpc_server_slots(sv2, 0, &m0);
  EXPECT(m0 > n0, "slots from free memory %lld", (long long)m0);
  fecl_ldpc_server_destroy(sv2);
  fake_set_device_budget(SIZE_MAX);
  fecl_ldpc_server_destroy(sv);
  EXPECT(fake_device_bytes_in_use() == before, "leak %zu", fake_device_bytes_in_use() - before);
}

static void test_oom_at_create() {
  auto a = make_handle(96, 48, 288, 6, FECL_MSG_F32);
  fecl_ldpc* hs[1] = {a.get()};
  const int64_t slots[1] = {64 * 40};
  const size_t before = fake_device_bytes_in_use();
  fake_set_device_budget(before + (200u << 10));   // less than 2560 slots need
  fecl_ldpc_server* sv = nullptr;
  const int rc = fecl_ldpc_server_create(hs, 1, slots, 0, &sv);
  EXPECT(rc == FECL_OUT_OF_MEMORY && sv == nullptr, "rc %d", rc);
  EXPECT(std::string(fecl_last_error()).find("out of memory") != std::string::npos, "message: %s",
         fecl_last_error());
  fake_set_device_budget(SIZE_MAX);
  EXPECT(fake_device_bytes_in_use() == before, "leak %zu", fake_device_bytes_in_use() - before);
}

static void test_oom_while_running() {
  auto a = make_handle(96, 48, 288, 30, FECL_MSG_F32);
  fecl_ldpc* hs[1] = {a.get()};
  const int64_t slots[1] = {64 * 20};
  fecl_ldpc_server* sv = nullptr;
  int rc = fecl_ldpc_server_create(hs, 1, slots, 0, &sv);
  EXPECT(rc == FECL_OK, "create: %s", fecl_last_error());
  if (rc) return;
  // staging holds one round's host frames (128 per step); 8 steps per round admit up to 1024
  // frames at once and must grow it, which the device can no longer afford
  fake_set_device_budget(fake_device_bytes_in_use() + 1024);
  fecl_ldpc_server_set_round_steps(sv, 8);
  std::mt19937_64 rng(7);
  Req r;
  r.frames = 600;
  fill_request(&r, a.get(), rng);
  rc = fecl_ldpc_server_submit(sv, 0, r.buf, r.frames, r.msg.data(), r.iters.data(), &r.ticket);
  EXPECT(rc == FECL_OK, "submit: %s", fecl_last_error());
  int32_t done = 1;
  rc = fecl_ldpc_server_wait(sv, r.ticket, 30.0, &done, nullptr);
  EXPECT(rc == FECL_OUT_OF_MEMORY && done == 0, "wait rc %d done %d (%s)", rc, done, fecl_last_error());
  std::printf("out of memory while running: %s\n", fecl_last_error());
  uint64_t t2 = 0;
  rc = fecl_ldpc_server_submit(sv, 0, r.buf, 1, r.msg.data(), nullptr, &t2);
  EXPECT(rc == FECL_OUT_OF_MEMORY, "submit after the failure: rc %d", rc);
  fake_set_device_budget(SIZE_MAX);
  fecl_ldpc_server_destroy(sv);
}

static void test_destroy_with_queued_requests() {
  auto a = make_handle(96, 48, 288, 40, FECL_MSG_F32);
  fecl_ldpc* hs[1] = {a.get()};
  const int64_t slots[1] = {64};
  fecl_ldpc_server* sv = nullptr;
  if (fecl_ldpc_server_create(hs, 1, slots, 0, &sv) != FECL_OK) {
    EXPECT(false, "create: %s", fecl_last_error());
    return;
  }
  std::mt19937_64 rng(11);
  std::vector<Req> rs(20);
  for (Req& r : rs) {
    r.frames = 50;
    fill_request(&r, a.get(), rng);
    EXPECT(fecl_ldpc_server_submit(sv, 0, r.buf, r.frames, r.msg.data(), r.iters.data(), &r.ticket) == FECL_OK,
           "submit");
  }
  fecl_ldpc_server_destroy(sv);   // must not hang or touch freed requests
}

int main() {
  test_mixed_requests(1);
  test_mixed_requests(3);
  test_admission_by_memory();
  test_oom_at_create();
  test_oom_while_running();
  test_destroy_with_queued_requests();
  std::printf("%d failures\n", g_failures);
  return g_failures ? 1 : 0;
}
