// CPU fake of the engine functions the serving loop calls (declared in ldpc_engine.h and
// ldpc_launch.h), for building fecl_amd/csrc/ldpc_server.cpp without a GPU.  Buffers come from
// the fake runtime's device memory; "kernels" are CPU functions enqueued on their stream, with
// the slot-level semantics the server relies on:
//   admit    copies a frame's channel values into its slot and sets its phase
//   step     check and variable updates do nothing; the control advances every live slot by
//            one iteration and retires it (PH_DONE) after fake_iterations() of them
//   extract  gives a retired slot's hard bits (llr < 0) and iteration count
//   copy_bytes  copies the byte ranges (the phase snapshot)
// Channel values sit slot-major ([slot][N]) here; nothing but these functions reads them.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "fake_engine.h"
#include "ldpc_engine.h"

namespace fecl {
namespace engine {

namespace {
thread_local std::string g_err;
}

int fail(int status, const std::string& msg) {
  g_err = msg;
  return status;
}

Layout layout_of(const fecl_ldpc* h) {
  Layout lay;
  lay.fmt = h->msg_format;
  lay.words = 1;
  lay.vpw = 1;
  return lay;
}

int copy_opts(const fecl_ldpc* h, OptsCopy* out) {
  out->opts = DevOpts{};
  out->opts.iterations = h->iterations;
  return FECL_OK;
}

int32_t nt_rows_for(const fecl_ldpc*, const Layout&, int64_t) { return 0; }

int prepare(fecl_ldpc* h, Workspace* ws, const Layout& lay, int64_t frames, bool, DevState* sp, DevOut* op) {
  const Code& c = h->code;
  const int64_t tw = lay.tw(), tiles = (frames + tw - 1) / tw, slots = tiles * tw;
  // the real engine's footprint per slot (messages, channel values, hard decisions, state)
  HIP_TRY(ws->msg.ensure((size_t)(slots * c.E * 4 / lay.pw())));
  HIP_TRY(ws->L.ensure((size_t)(slots * c.N * 4)));
  HIP_TRY(ws->hard.ensure((size_t)(tiles * c.N * lay.fpl() * 8)));
  HIP_TRY(ws->phase.ensure((size_t)slots));
  HIP_TRY(ws->iter.ensure((size_t)slots * 4));
  HIP_TRY(ws->fail.ensure((size_t)(tiles * lay.fpl() * 8)));
  HIP_TRY(ws->flags.ensure((size_t)tiles * 8));
  HIP_TRY(ws->out_hard.ensure((size_t)(tiles * c.K * lay.fpl() * 8)));
  DevState s{};
  s.msg = ws->msg.as<uint32_t>();
  s.L = ws->L.as<float>();
  s.hard = ws->hard.as<uint64_t>();
  s.phase = ws->phase.as<uint8_t>();
  s.iter = ws->iter.as<int32_t>();
  s.fail = ws->fail.as<uint64_t>();
  s.tile_flags = ws->flags.as<uint32_t>();
  s.tile_iter = ws->flags.as<int32_t>() + tiles;
  s.active = nullptr;
  s.tiles = (int32_t)tiles;
  *sp = s;
  DevOut o{};
  o.hard = ws->out_hard.as<uint64_t>();
  *op = o;
  return FECL_OK;
}

}  // namespace engine

hipError_t kernels_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st) {
  fake_enqueue(st, [=] {
    for (int64_t i = 0; i < n; ++i) p[i] = v;
  });
  return hipSuccess;
}

hipError_t kernels_copy_bytes(const ByteCopies& c, hipStream_t st) {
  if (c.count < 0 || c.count > kMaxByteCopies) return hipErrorInvalidValue;
  fake_enqueue(st, [c] {
    for (int i = 0; i < c.count; ++i) std::memcpy(c.item[i].dst, c.item[i].src, (size_t)c.item[i].n);
  });
  return hipSuccess;
}

hipError_t kernels_admit(const Layout&, const DevCode& c, const DevState& s, const float* stage, const int32_t* slots,
                         int n, uint8_t init_phase, hipStream_t st) {
  fake_enqueue(st, [=] {
    for (int i = 0; i < n; ++i) {
      const int32_t slot = slots[i];
      std::memcpy(s.L + (int64_t)slot * c.N, stage + (int64_t)i * c.N, (size_t)c.N * 4);
      s.phase[slot] = init_phase;
      s.iter[slot] = 0;
    }
  });
  return hipSuccess;
}

hipError_t kernels_check(const Layout&, const DevCode&, const DevState&, const DevOpts&, int, hipStream_t st) {
  fake_enqueue(st, [] {});
  return hipSuccess;
}

hipError_t kernels_variable(const Layout&, const DevCode&, const DevState&, const DevOut&, int, int, bool,
                            hipStream_t st) {
  fake_enqueue(st, [] {});
  return hipSuccess;
}

hipError_t kernels_syndrome_control(const Layout& lay, const DevCode& c, const DevState& s, int iterations, int, int,
                                    int, hipStream_t st) {
  const int64_t slots = (int64_t)s.tiles * lay.tw();
  fake_enqueue(st, [=] {
    for (int64_t slot = 0; slot < slots; ++slot) {
      const int base = s.phase[slot] & PH_MASK;
      if (base != PH_RUN && base != PH_FINAL) continue;
      const int it = ++s.iter[slot];
      if (it >= fake_iterations(s.L + slot * c.N, iterations)) s.phase[slot] = PH_DONE;
      else s.phase[slot] = PH_RUN;
    }
  });
  return hipSuccess;
}

hipError_t kernels_extract(const Layout&, const DevCode& c, const DevState& s, const DevOut&, const int32_t* slots,
                           int n, uint8_t* bits, int32_t* iters, hipStream_t st) {
  fake_enqueue(st, [=] {
    for (int i = 0; i < n; ++i) {
      const int32_t slot = slots[i];
      const float* l = s.L + (int64_t)slot * c.N;
      for (int32_t v = 0; v < c.K; ++v) bits[(int64_t)i * c.K + v] = l[v] < 0.0f ? 1 : 0;
      iters[i] = s.iter[slot];
    }
  });
  return hipSuccess;
}

}  // namespace fecl

extern "C" const char* fecl_last_error(void) { return fecl::engine::g_err.c_str(); }
