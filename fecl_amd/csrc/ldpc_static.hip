// Decode that allocates nothing (fecl_ldpc_decode_device_noalloc): the batch's device
// workspace lives in `scratch`, a fixed array compiled into the library, so the call itself
// makes no hipMalloc/hipFree/hipHostMalloc/event or stream creation and may be recorded into
// a hipGraph or run where allocation is forbidden.  It decodes like fecl_ldpc_decode_device
// (same kernels, same results) with three differences:
//   - one lane word per edge (64-frame tiles for FP32 messages, 128 for 16-bit), so small
//     batches fit; the layout does not change results;
//   - every one of the `iterations` steps is enqueued (no host reads the live count), as in a
//     captured graph: steps after the last frame finishes return at once;
//   - no compaction and no passes: the whole batch must fit in `scratch`.
// Every call on a device shares that device's array and leaves its contents behind; the
// ingest kernel resets all per-frame state a call reads, so nothing carries over.  Calls that
// overlap on one device (different streams, or different host threads without ordering) race
// on the array.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "ldpc_engine.h"

namespace fecl {
namespace engine {

// Capacity of the no-allocation workspace (per device, held from the first call until the
// process ends).  It fits one 64-frame FP32 tile or one 128-frame FP16/BF16 tile of DVB-S2
// n = 64800 rate 1/2 (75 / 93 MB), or several tiles of short frames.
constexpr size_t kScratchBytes = size_t(128) << 20;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;   // per device: the device address of `scratch`

// The device address of `scratch` on `device` (current), looked up once per device.
int scratch_address(int device, uint8_t** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0) return fail(FECL_DEVICE_ERROR, "invalid device");
  if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[(size_t)device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(FECL_DEVICE_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[(size_t)device] = p;
  }
  *out = static_cast<uint8_t*>(g_scratch_ptr[(size_t)device]);
  return FECL_OK;
}

// Bump allocation of 256-byte aligned pieces of the workspace; with a null base it only
// counts the bytes.
struct Carve {
  uint8_t* base;
  size_t off = 0;
  template <typename T> T* take(size_t bytes) {
    const size_t at = (off + 255) & ~size_t(255);
    off = at + bytes;
    return base ? reinterpret_cast<T*>(base + at) : nullptr;
  }
};

// Device views of a `tiles`-tile region of the layout, carved from c (decode outputs only).
void carve_region(Carve* c, const Code& code, const Layout& lay, int64_t tiles, DevState* s, DevOut* out) {
  const int FPL = lay.fpl(), TW = lay.tw();
  *s = DevState{};
  *out = DevOut{};
  s->msg = c->take<uint32_t>((size_t)tiles * code.E * lay.rw() * 4);
  s->L = c->take<float>((size_t)tiles * code.N * TW * 4);
  s->hard = c->take<uint64_t>((size_t)tiles * code.N * FPL * 8);
  s->phase = c->take<uint8_t>((size_t)tiles * TW);
  s->iter = c->take<int32_t>((size_t)tiles * TW * 4);
  s->fail = c->take<uint64_t>((size_t)tiles * FPL * 8);
  s->tile_flags = c->take<uint32_t>((size_t)tiles * 4);
  s->tile_iter = c->take<int32_t>((size_t)tiles * 4);
  s->active = nullptr;   // the control kernel counts live frames only for a step >= 0
  s->tiles = (int32_t)tiles;
  out->hard = c->take<uint64_t>((size_t)tiles * code.K * FPL * 8);
}

}  // namespace

int decode_noalloc(fecl_ldpc* h, const float* llr, int64_t frames, uint8_t* msg, int32_t* iterations,
                   hipStream_t st) {
  if (int rc = check_handle(h)) return rc;
  if (frames < 0) return fail(FECL_INVALID_ARGUMENT, "Invalid size for parity");
  if (frames == 0) return FECL_OK;
  if (!llr || !msg) return fail(FECL_INVALID_ARGUMENT, "Input must contains parity");
  HIP_TRY(hipSetDevice(h->device));
  uint8_t* ptr = nullptr;
  if (int rc = scratch_address(h->device, &ptr)) return rc;
  const Code& c = h->code;
  Layout lay;
  lay.fmt = h->msg_format;
  lay.words = 1;
  lay.vpw = h->vpw;
  if (!layout_fits(lay, h->algorithm, c.max_row_degree, c.max_var_degree))
    return fail(FECL_NOT_SUPPORTED, "no register-budget layout with one word per lane for this code");
  if (!vpw_fits(lay, c.max_var_degree)) lay.vpw = 1;
  const int64_t tiles = (frames + lay.tw() - 1) / lay.tw();
  DevState s;
  DevOut out;
  Carve need{nullptr};
  carve_region(&need, c, lay, tiles, &s, &out);
  if (need.off > kScratchBytes)
    return fail(FECL_INVALID_ARGUMENT, "fecl_ldpc_decode_device_noalloc: " + std::to_string(frames) +
                                           " frames of this code need " + std::to_string(need.off) +
                                           " bytes of workspace; the library holds " +
                                           std::to_string(kScratchBytes));
  Carve carve{ptr};
  carve_region(&carve, c, lay, tiles, &s, &out);
  DevOpts o = make_opts(h);
  o.nt_rows = 0;
  const uint8_t init = (uint8_t)((h->iterations == 1 ? PH_FINAL : PH_RUN) | PH_FRESH);
  HIP_TRY(kernels_ingest(lay, h->dc, s, llr, nullptr, false, frames, 0, init, st));
  for (int step = 0; step < h->iterations; ++step) {
    HIP_TRY(kernels_check(lay, h->dc, s, o, c.max_row_degree, st));
    HIP_TRY(kernels_variable(lay, h->dc, s, out, c.max_var_degree, o.fused_syn, false, st));
    HIP_TRY(kernels_syndrome_control(lay, h->dc, s, h->iterations, o.fused_syn, -1, c.max_row_degree, st));
  }
  const BatchResults res{msg, iterations, nullptr, nullptr, nullptr, nullptr};
  HIP_TRY(kernels_results(lay, h->dc, s, out, frames, 0, res, false, nullptr, nullptr, st));
  return FECL_OK;
}

}  // namespace engine
}  // namespace fecl

extern "C" int fecl_ldpc_decode_device_noalloc(fecl_ldpc* h, const float* llr, int64_t frames, uint8_t* msg,
                                               int32_t* iterations, void* stream) {
  return fecl::engine::decode_noalloc(h, llr, frames, msg, iterations, (hipStream_t)stream);
}
