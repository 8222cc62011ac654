// Host entry points into the kernels (ldpc_kernels.hip).  fmt = fecl_msg_format.
#pragma once
#include <hip/hip_runtime.h>

#include "ldpc_kernels.h"

namespace fecl {

// Where the results of a batch go (device pointers, frame-major; null = not wanted).
// msg/syst/parity/state are double when `dbl`, float otherwise.
struct BatchResults {
  uint8_t* msg_bits;     // [F][K] 0/1
  int32_t* iterations;   // [F]
  void* msg;             // [F][K]  a-posteriori (soDecode)
  void* syst;            // [F][K]  extrinsic
  void* parity;          // [F][N]  extrinsic
  void* state;           // [F][E]  check-to-variable
};

// Device layout of a workspace: message format and the 32-bit words each lane owns per edge.
struct Layout {
  int fmt = 0;     // fecl_msg_format
  int words = 1;   // 1, 2 or 4
  int vpw = 1;     // variables per wavefront in the variable kernel: 1 or 2
  int pw() const { return fmt == 0 ? 1 : 2; }        // messages per word
  int fpl() const { return words * pw(); }           // frames per lane
  int tw() const { return 64 * fpl(); }              // frames per tile
  int rw() const { return 64 * words; }              // words per (tile, edge) row
};

// Whether the kernels for this layout stay within the register budget for the code's degrees.
bool layout_fits(const Layout& lay, int algorithm, int max_dc, int max_dv);

// Whether the check kernel for this layout holds a row in registers (rather than streaming it
// twice from memory, the fallback for long rows).
bool cn_register_staged(const Layout& lay, int algorithm, int sf_mode, int max_dc);

// Whether vpw variables per wavefront fit the variable kernel's register budget.
bool vpw_fits(const Layout& lay, int max_dv);

// Degree bound of the variable kernel for a code (the stride of DevCode::var_tab).
inline int var_table_stride(int max_dv) { return max_dv <= 4 ? 4 : max_dv <= 8 ? 8 : max_dv <= 16 ? 16 : 32; }

hipError_t kernels_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st);
// Byte ranges copied by one kernel launch, in stream order (the serving loop snapshots every
// pool's frame phases into mapped page-locked memory with it).
struct ByteCopy {
  const uint8_t* src;
  uint8_t* dst;
  int64_t n;
};
constexpr int kMaxByteCopies = 32;
struct ByteCopies {
  ByteCopy item[kMaxByteCopies];
  int32_t count;
};
hipError_t kernels_copy_bytes(const ByteCopies& c, hipStream_t st);
// Permutation of FP32 blocks: out[f][i] = in[f][seq[i]] (permute), or out[f][seq[i]] = in[f][i]
// with the other entries zeroed (inverse).
hipError_t kernels_permute(const int64_t* seq, int64_t out_size, int64_t in_size, const float* in, float* out,
                           int64_t frames, bool inverse, hipStream_t st);
// Systematic staircase encoder: msg [frames][K] (0/1 bytes) -> codeword [frames][N].
hipError_t kernels_encode(const DevCode& c, const uint8_t* msg, uint8_t* cw, int64_t frames, hipStream_t st);
hipError_t kernels_ingest(const Layout& lay, const DevCode& c, const DevState& s, const void* llr, const void* syst,
                          bool dbl, int64_t frames, int64_t frame0, uint8_t init_phase, hipStream_t st);
hipError_t kernels_state_in(const Layout& lay, const DevCode& c, const DevState& s, const void* state, bool dbl,
                            int64_t frames, int64_t frame0, hipStream_t st);
hipError_t kernels_vn_init(const Layout& lay, const DevCode& c, const DevState& s, int max_dv, hipStream_t st);
hipError_t kernels_check(const Layout& lay, const DevCode& c, const DevState& s, const DevOpts& o, int max_dc,
                         hipStream_t st);
// nt: non-temporal loads and stores for the message rows (see DevOpts::nt_rows)
hipError_t kernels_variable(const Layout& lay, const DevCode& c, const DevState& s, const DevOut& out, int max_dv,
                            int fused, bool nt, hipStream_t st);
hipError_t kernels_syndrome_control(const Layout& lay, const DevCode& c, const DevState& s, int iterations, int fused,
                                    int step, int max_dc, hipStream_t st);
// Serving: write admitted frames' LLRs into their slots; read finished slots' results.
hipError_t kernels_admit(const Layout& lay, const DevCode& c, const DevState& s, const float* stage,
                         const int32_t* slots, int n, uint8_t init_phase, hipStream_t st);
hipError_t kernels_extract(const Layout& lay, const DevCode& c, const DevState& s, const DevOut& out,
                           const int32_t* slots, int n, uint8_t* bits, int32_t* iters, hipStream_t st);
// Results of a region of tiles; map[j] is the frame of slot j (-1: none), or null for the
// batch's first region, whose slot j holds frame j.
// ctl (nullable): the results are written only if ctl->go.
hipError_t kernels_results(const Layout& lay, const DevCode& c, const DevState& s, const DevOut& out, int64_t frames,
                           int64_t frame0, const BatchResults& r, bool dbl, const int32_t* map, const CompactCtl* ctl,
                           hipStream_t st);
// Moves the live frames of region `from` (slot j holds frame map_in[j], or j when map_in is
// null) in slot order into the first slots of region `to`, which must have room for them:
// message rows, channel values, phases and iteration counts; sets map_out (the frames of
// `to`'s slots, -1 past the last) and `to`'s tile flags.  src_of: scratch of to.tiles * TW
// entries.  Call at a step boundary (after the control kernel), without the fused syndrome.
hipError_t kernels_compact(const Layout& lay, const DevCode& c, const DevState& from, const DevState& to,
                           const int32_t* map_in, int32_t* map_out, int32_t* src_of, hipStream_t st);
// Compaction decided on the device (captured graphs), in place in one region: the plan sets
// ctl (go, live) from the phases, the steps left (`remaining`) and the tiles holding live
// frames, and writes map_out (a copy of map_in when it declines) and src_of (s.tiles * TW
// entries each).  Between the plan and the move the caller hands the region's finished
// frames' results over with the same ctl.  The move (one launch, nothing unless ctl->go)
// takes the live frames into s's first slots in one pass and resets s's tile flags.
hipError_t kernels_compact_plan(const Layout& lay, const DevState& s, const int32_t* map_in, int32_t* map_out,
                                int32_t* src_of, CompactCtl* ctl, int32_t remaining, hipStream_t st);
hipError_t kernels_compact_in_place(const Layout& lay, const DevCode& c, const DevState& s, const int32_t* src_of,
                                    const CompactCtl* ctl, hipStream_t st);

}  // namespace fecl
