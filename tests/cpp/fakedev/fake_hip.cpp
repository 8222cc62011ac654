// CPU test double of the HIP runtime (see hip/hip_runtime.h).
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

struct FakeStream {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::function<void()>> q;
  uint64_t enqueued = 0, done = 0;
  bool stop = false;
  std::thread th;
  void run() {
    for (;;) {
      std::function<void()> fn;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !q.empty(); });
        if (q.empty()) return;
        fn = std::move(q.front());
        q.pop_front();
      }
      fn();
      {
        std::lock_guard<std::mutex> lk(mu);
        ++done;
      }
      cv.notify_all();
    }
  }
  void wait_for(uint64_t pos) {
    std::unique_lock<std::mutex> lk(mu);
    cv.wait(lk, [&] { return done >= pos; });
  }
};

struct FakeEvent {
  std::mutex mu;
  FakeStream* s = nullptr;   // null: never recorded, or recorded on the null stream (complete)
  uint64_t pos = 0;
};

namespace {

std::mutex g_mem_mu;
std::map<const void*, size_t> g_dev, g_host;   // live allocations
size_t g_budget = SIZE_MAX, g_in_use = 0;

bool find(const std::map<const void*, size_t>& m, const void* p) {
  auto it = m.upper_bound(p);
  if (it == m.begin()) return false;
  --it;
  return (const char*)p < (const char*)it->first + it->second;
}

}  // namespace

void fake_set_device_budget(size_t bytes) {
  std::lock_guard<std::mutex> lk(g_mem_mu);
  g_budget = bytes;
}
size_t fake_device_bytes_in_use(void) {
  std::lock_guard<std::mutex> lk(g_mem_mu);
  return g_in_use;
}

void fake_enqueue(hipStream_t st, std::function<void()> fn) {
  if (!st) {
    fn();
    return;
  }
  {
    std::lock_guard<std::mutex> lk(st->mu);
    st->q.push_back(std::move(fn));
    ++st->enqueued;
  }
  st->cv.notify_all();
}

hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetLastError(void) { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) {
  switch (e) {
    case hipSuccess: return "no error";
    case hipErrorOutOfMemory: return "out of memory";
    case hipErrorInvalidValue: return "invalid argument";
    default: return "unknown error";
  }
}

hipError_t hipMalloc(void** p, size_t bytes) {
  if (!p) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lk(g_mem_mu);
  if (bytes > g_budget - g_in_use) return hipErrorOutOfMemory;
  void* m = std::calloc(1, bytes ? bytes : 1);
  if (!m) return hipErrorOutOfMemory;
  g_dev[m] = bytes;
  g_in_use += bytes;
  *p = m;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  std::lock_guard<std::mutex> lk(g_mem_mu);
  auto it = g_dev.find(p);
  if (it == g_dev.end()) return hipErrorInvalidValue;
  g_in_use -= it->second;
  g_dev.erase(it);
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
  if (!p) return hipErrorInvalidValue;
  void* m = std::calloc(1, bytes ? bytes : 1);
  if (!m) return hipErrorOutOfMemory;
  std::lock_guard<std::mutex> lk(g_mem_mu);
  g_host[m] = bytes;
  *p = m;
  return hipSuccess;
}
hipError_t hipHostFree(void* p) {
  if (!p) return hipSuccess;
  std::lock_guard<std::mutex> lk(g_mem_mu);
  auto it = g_host.find(p);
  if (it == g_host.end()) return hipErrorInvalidValue;
  g_host.erase(it);
  std::free(p);
  return hipSuccess;
}
hipError_t hipHostGetDevicePointer(void** dev, void* host, unsigned) {
  *dev = host;   // mapped: the same memory
  return hipSuccess;
}
hipError_t hipMemGetInfo(size_t* free_bytes, size_t* total_bytes) {
  std::lock_guard<std::mutex> lk(g_mem_mu);
  const size_t total = g_budget == SIZE_MAX ? (size_t(1) << 30) : g_budget;
  if (free_bytes) *free_bytes = total > g_in_use ? total - g_in_use : 0;
  if (total_bytes) *total_bytes = total;
  return hipSuccess;
}
hipError_t hipPointerGetAttributes(hipPointerAttribute_t* attr, const void* p) {
  std::lock_guard<std::mutex> lk(g_mem_mu);
  std::memset(attr, 0, sizeof(*attr));
  if (find(g_dev, p)) attr->type = hipMemoryTypeDevice;
  else if (find(g_host, p)) attr->type = hipMemoryTypeHost;
  else return hipErrorInvalidValue;   // pageable memory is unknown to the runtime
  attr->devicePointer = attr->hostPointer = const_cast<void*>(p);
  return hipSuccess;
}

hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
  hipDeviceSynchronize();
  std::memcpy(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t st) {
  fake_enqueue(st, [=] { std::memcpy(dst, src, bytes); });
  return hipSuccess;
}
hipError_t hipMemset(void* p, int value, size_t bytes) {
  hipDeviceSynchronize();
  std::memset(p, value, bytes);
  return hipSuccess;
}

namespace {
std::mutex g_streams_mu;
std::map<FakeStream*, bool> g_streams;
}  // namespace

hipError_t hipStreamCreateWithFlags(hipStream_t* st, unsigned) {
  FakeStream* s = new FakeStream();
  s->th = std::thread([s] { s->run(); });
  std::lock_guard<std::mutex> lk(g_streams_mu);
  g_streams[s] = true;
  *st = s;
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t st) {
  if (!st) return hipSuccess;
  {
    std::lock_guard<std::mutex> lk(g_streams_mu);
    g_streams.erase(st);
  }
  {
    std::lock_guard<std::mutex> lk(st->mu);
    st->stop = true;
  }
  st->cv.notify_all();
  st->th.join();   // runs what is still queued first
  delete st;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t st) {
  if (!st) return hipSuccess;
  uint64_t pos;
  {
    std::lock_guard<std::mutex> lk(st->mu);
    pos = st->enqueued;
  }
  st->wait_for(pos);
  return hipSuccess;
}
hipError_t hipDeviceSynchronize(void) {
  std::vector<FakeStream*> all;
  {
    std::lock_guard<std::mutex> lk(g_streams_mu);
    for (auto& kv : g_streams) all.push_back(kv.first);
  }
  for (FakeStream* s : all) hipStreamSynchronize(s);
  return hipSuccess;
}

hipError_t hipEventCreateWithFlags(hipEvent_t* ev, unsigned) {
  *ev = new FakeEvent();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t ev) {
  delete ev;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t ev, hipStream_t st) {
  std::lock_guard<std::mutex> lk(ev->mu);
  ev->s = st;
  if (st) {
    std::lock_guard<std::mutex> sl(st->mu);
    ev->pos = st->enqueued;
  } else {
    ev->pos = 0;
  }
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t ev) {
  FakeStream* s;
  uint64_t pos;
  {
    std::lock_guard<std::mutex> lk(ev->mu);
    s = ev->s;
    pos = ev->pos;
  }
  if (s) s->wait_for(pos);
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t st, hipEvent_t ev, unsigned) {
  FakeStream* s;
  uint64_t pos;
  {
    std::lock_guard<std::mutex> lk(ev->mu);
    s = ev->s;
    pos = ev->pos;
  }
  if (!s || s == st) return hipSuccess;   // complete, or already ordered by the stream
  fake_enqueue(st, [s, pos] { s->wait_for(pos); });
  return hipSuccess;
}
