// Host-side conversions of the host-buffer decode path (ldpc_capi.cpp): FP64 inputs rounded to
// FP32 and FP32 outputs widened to FP64, split over a few threads that concurrent callers share.
// Pure C++, so the CPU tests run it under ThreadSanitizer.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <thread>
#include <vector>

namespace fecl {
namespace engine {

constexpr unsigned kConvertThreads = 8;

// Runs body(a, b) over [0, n) split across up to kConvertThreads threads.
// Host-buffer calls in flight in this process: FeCl's fan-out calls the hooks from
// workGroupSize threads at once, and the conversion threads are shared out among them.
inline std::atomic<int> g_host_calls{0};

template <class Body>
void parallel_ranges(size_t n, Body body) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t share = std::max<size_t>(1, kConvertThreads / (size_t)std::max(1, g_host_calls.load()));
  const size_t threads = std::max<size_t>(1, std::min<size_t>({share, hw, n / 65536 + 1}));
  const size_t per = (n + threads - 1) / threads;
  auto part = [&](size_t t) { body(std::min(n, per * t), std::min(n, per * (t + 1))); };
  std::vector<std::thread> pool;
  for (size_t t = 1; t < threads; ++t) pool.emplace_back(part, t);
  part(0);
  for (auto& th : pool) th.join();
}

// dst[i] = (float)src[i]: round to nearest even, as the device's double-to-float conversion.
inline void to_float(const double* src, float* dst, size_t n) {
  parallel_ranges(n, [=](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) dst[i] = (float)src[i];
  });
}

// A-priori inputs on the first K bits (soDecode): dst[f][v] = (float)((double)(float)llr[f][v]
// + syst[f][v]) for v < K, (float)llr[f][v] otherwise.  FeCl's decoder instantiated with a
// float LLR type stores the channel value as float, then adds the double a-priori value and
// rounds once (src/detail/BpDecoder/BpDecoderImpl.cpp:75-80); rounding syst to float first
// would round twice.
inline void to_float_plus(const double* llr, const double* syst, float* dst, int64_t frames, int64_t N, int64_t K) {
  parallel_ranges((size_t)frames * N, [=](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) {
      const size_t f = i / (size_t)N, v = i % (size_t)N;
      const float x = (float)llr[i];
      dst[i] = v < (size_t)K ? (float)((double)x + syst[f * (size_t)K + v]) : x;
    }
  });
}

inline void to_double(const float* src, double* dst, size_t n) {
  parallel_ranges(n, [=](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) dst[i] = (double)src[i];
  });
}

inline void copy_bytes(const void* src, void* dst, size_t n) {
  parallel_ranges(n, [=](size_t a, size_t b) { std::memcpy((char*)dst + a, (const char*)src + a, b - a); });
}

}  // namespace engine
}  // namespace fecl
