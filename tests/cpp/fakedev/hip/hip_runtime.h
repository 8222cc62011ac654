// CPU test double of the HIP runtime calls the serving loop (fecl_amd/csrc/ldpc_server.cpp)
// makes, for building that file with g++ and running it without a GPU (and under ASan/TSan).
//
// "Device" memory is host memory from malloc, charged against a budget that a test sets
// (fake_set_device_budget), so device out-of-memory can be forced at any allocation.  Streams
// are real asynchronous queues: each runs its operations in order on a thread of its own, and
// events mark positions in a stream, so host code that reads a result before synchronizing on
// it races exactly as it would on a GPU (and TSan sees it).  The null stream runs operations
// at once.  Kernel launches are the fake engine's CPU functions (fake_engine.cpp), enqueued with
// fake_enqueue.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <functional>

typedef enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNotReady = 600,
} hipError_t;

typedef struct FakeStream* hipStream_t;
typedef struct FakeEvent* hipEvent_t;

enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum hipMemoryType { hipMemoryTypeUnregistered = 0, hipMemoryTypeHost = 1, hipMemoryTypeDevice = 2 };
struct hipPointerAttribute_t {
  hipMemoryType type;
  int device;
  void* devicePointer;
  void* hostPointer;
};

#define hipStreamNonBlocking 0x1u
#define hipEventDisableTiming 0x2u
#define hipHostMallocDefault 0x0u
#define hipHostMallocMapped 0x2u

hipError_t hipSetDevice(int device);
hipError_t hipGetLastError(void);
const char* hipGetErrorString(hipError_t e);
hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipHostGetDevicePointer(void** dev, void* host, unsigned flags);
hipError_t hipMemGetInfo(size_t* free_bytes, size_t* total_bytes);
hipError_t hipPointerGetAttributes(hipPointerAttribute_t* attr, const void* p);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t st);
hipError_t hipMemset(void* p, int value, size_t bytes);
hipError_t hipStreamCreateWithFlags(hipStream_t* st, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t st);
hipError_t hipStreamSynchronize(hipStream_t st);
hipError_t hipStreamWaitEvent(hipStream_t st, hipEvent_t ev, unsigned flags);
hipError_t hipEventCreateWithFlags(hipEvent_t* ev, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t ev);
hipError_t hipEventRecord(hipEvent_t ev, hipStream_t st);
hipError_t hipEventSynchronize(hipEvent_t ev);
hipError_t hipDeviceSynchronize(void);

// ---- test controls
// Device bytes that may be allocated in all (default: unlimited).
void fake_set_device_budget(size_t bytes);
size_t fake_device_bytes_in_use(void);
// Runs fn on stream st after everything enqueued before it (at once on the null stream).
void fake_enqueue(hipStream_t st, std::function<void()> fn);
