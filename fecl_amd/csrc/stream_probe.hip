// Achievable-bandwidth probe for the benchmark (libfecl_stream.so; not part of the decoder):
// streaming kernels over 16-byte words that copy (read and write), only read, or only write,
// with or without non-temporal hints, in one of two orders: grid-stride (four loads in flight
// per thread, each wave instruction covering 1 KiB, consecutive instructions a whole grid
// apart) or per-block chunks (each workgroup walks its own contiguous chunk, eight loads in
// flight per thread); and an in-place update of scattered 1 KiB rows.  The copy and the
// update move 2 x bytes, the others bytes.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ inline u32x4 ld(const u32x4* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  return *p;
}
template <bool NT>
__device__ inline void stv(u32x4* p, u32x4 v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <bool NT>
__global__ __launch_bounds__(256) void copy_kernel(const u32x4* src, u32x4* dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    const u32x4 a = ld<NT>(src + i), b = ld<NT>(src + i + stride), c = ld<NT>(src + i + 2 * stride),
                d = ld<NT>(src + i + 3 * stride);
    stv<NT>(dst + i, a);
    stv<NT>(dst + i + stride, b);
    stv<NT>(dst + i + 2 * stride, c);
    stv<NT>(dst + i + 3 * stride, d);
  }
  for (; i < n; i += stride) stv<NT>(dst + i, ld<NT>(src + i));
}

template <bool NT>
__global__ __launch_bounds__(256) void read_kernel(const u32x4* src, u32x4* sink, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  u32x4 acc = {0, 0, 0, 0};
  for (; i + 3 * stride < n; i += 4 * stride) {
    const u32x4 a = ld<NT>(src + i), b = ld<NT>(src + i + stride), c = ld<NT>(src + i + 2 * stride),
                d = ld<NT>(src + i + 3 * stride);
    acc ^= a ^ b ^ c ^ d;
  }
  for (; i < n; i += stride) acc ^= ld<NT>(src + i);
  // one word per thread keeps the loads live
  sink[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

template <bool NT>
__global__ __launch_bounds__(256) void write_kernel(u32x4* dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const u32x4 v = {(unsigned)threadIdx.x, 1u, 2u, 3u};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) stv<NT>(dst + i, v);
}

// Scattered rows, in place: each wavefront reads and writes back eight 1 KiB rows per round
// (one 16-byte word per lane per row), the rows picked by a hash of (round, wavefront, slot)
// over the whole buffer, as the variable pass gathers and updates a variable's edge rows.
// Rounds are sized so that the rows visited add up to the buffer's size.
__device__ inline uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

template <bool NT>
__global__ __launch_bounds__(256) void scatter_kernel(u32x4* buf, int64_t rows, int rounds) {
  const int lane = threadIdx.x & 63;
  const uint32_t wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint32_t waves = gridDim.x * 4;
  for (int r = 0; r < rounds; ++r) {
    int64_t off[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
      off[k] = (int64_t)(mix32(((uint32_t)r * waves + wave) * 8u + (uint32_t)k) % (uint64_t)rows) * 64 + lane;
    u32x4 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = ld<NT>(buf + off[k]);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      v[k].x += 1u;
      stv<NT>(buf + off[k], v[k]);
    }
  }
}

// Per-block chunks: workgroup b owns words [b * chunk, (b + 1) * chunk) (chunk a multiple
// of 256 * 8) and walks them 256 * 8 words (32 KiB) at a time.
template <bool NT, int MODE>
__global__ __launch_bounds__(256) void chunk_kernel(const u32x4* src, u32x4* dst, int64_t n, int64_t chunk) {
  constexpr int U = 8;
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  u32x4 acc = {0, 0, 0, 0};
  for (int64_t base = lo; base < hi; base += 256 * U) {
    const int64_t i0 = base + threadIdx.x;
    if (MODE == 2) {
      const u32x4 v = {(unsigned)threadIdx.x, 1u, 2u, 3u};
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (i0 + 256 * k < hi) stv<NT>(dst + i0 + 256 * k, v);
    } else if (base + 256 * U <= hi) {
      u32x4 v[U];
#pragma unroll
      for (int k = 0; k < U; ++k) v[k] = ld<NT>(src + i0 + 256 * k);
#pragma unroll
      for (int k = 0; k < U; ++k) {
        if (MODE == 0) stv<NT>(dst + i0 + 256 * k, v[k]);
        else acc ^= v[k];
      }
    } else {
      for (int k = 0; k < U; ++k) {
        const int64_t i = i0 + 256 * k;
        if (i < hi) {
          if (MODE == 0) stv<NT>(dst + i, ld<NT>(src + i));
          else acc ^= ld<NT>(src + i);
        }
      }
    }
  }
  if (MODE == 1) dst[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

}  // namespace

// mode 0: copy src -> dst; 1: read src (dst receives blocks * 256 words of xor sums, so it
// needs blocks * 4 KiB beside nothing else); 2: write dst; 3: scattered 1 KiB rows of dst
// updated in place (src unused; moves about 2 x bytes, exactly 2 x 8 KiB x rounds x waves).
// flags bit 0: non-temporal loads and stores; bit 1: per-block chunks instead of grid-stride
// order (modes 0-2).  bytes is a multiple of 16 (of 1024 for mode 3); blocks = 0 picks 8 per
// CU.  Returns 0, 1 for bad arguments, 3 for a HIP error.
extern "C" int fecl_stream_probe(int mode, const void* src, void* dst, int64_t bytes, int flags, int blocks,
                                 void* stream) {
  if (bytes <= 0 || bytes % 16 || mode < 0 || mode > 3 || flags < 0 || flags > 3) return 1;
  if (mode == 3 && (bytes % 1024 || bytes / 1024 > 0xffffffffLL)) return 1;
  if (blocks <= 0) {
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 3;
    blocks = 8 * prop.multiProcessorCount;
  }
  const int64_t n = bytes / 16;
  const bool nt = flags & 1;
  hipStream_t st = (hipStream_t)stream;
  const auto* s = static_cast<const u32x4*>(src);
  auto* d = static_cast<u32x4*>(dst);
  if (mode == 3) {
    const int64_t rows = bytes / 1024;
    const int64_t waves = (int64_t)blocks * 4;
    const int rounds = (int)std::max<int64_t>(1, rows / (8 * waves));
    if (nt) hipLaunchKernelGGL(scatter_kernel<true>, dim3(blocks), dim3(256), 0, st, d, rows, rounds);
    else hipLaunchKernelGGL(scatter_kernel<false>, dim3(blocks), dim3(256), 0, st, d, rows, rounds);
  } else if (flags & 2) {
    const int64_t step = 256 * 8;
    const int64_t chunk = ((n + blocks - 1) / blocks + step - 1) / step * step;
    const int nb = (int)((n + chunk - 1) / chunk);   // no workgroup with an empty chunk
#define FECL_CHUNK(NT_, M_) hipLaunchKernelGGL((chunk_kernel<NT_, M_>), dim3(nb), dim3(256), 0, st, s, d, n, chunk)
    if (mode == 0) { if (nt) FECL_CHUNK(true, 0); else FECL_CHUNK(false, 0); }
    else if (mode == 1) { if (nt) FECL_CHUNK(true, 1); else FECL_CHUNK(false, 1); }
    else { if (nt) FECL_CHUNK(true, 2); else FECL_CHUNK(false, 2); }
#undef FECL_CHUNK
  } else if (mode == 0) {
    if (nt) hipLaunchKernelGGL(copy_kernel<true>, dim3(blocks), dim3(256), 0, st, s, d, n);
    else hipLaunchKernelGGL(copy_kernel<false>, dim3(blocks), dim3(256), 0, st, s, d, n);
  } else if (mode == 1) {
    if (nt) hipLaunchKernelGGL(read_kernel<true>, dim3(blocks), dim3(256), 0, st, s, d, n);
    else hipLaunchKernelGGL(read_kernel<false>, dim3(blocks), dim3(256), 0, st, s, d, n);
  } else {
    if (nt) hipLaunchKernelGGL(write_kernel<true>, dim3(blocks), dim3(256), 0, st, d, n);
    else hipLaunchKernelGGL(write_kernel<false>, dim3(blocks), dim3(256), 0, st, d, n);
  }
  return hipGetLastError() == hipSuccess ? 0 : 3;
}
