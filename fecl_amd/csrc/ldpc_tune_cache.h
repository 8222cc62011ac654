// Autotune cache: the lane layout fecl_ldpc_autotune picked, kept per (GPU architecture and
// compute-unit count, code, algorithm, message format, batch-size bucket) so that later
// autotunes of the same key pin it without probing.  Pure C++ (no device code),
// so the CPU tests and the sanitizer builds exercise it directly.
//
// The optional file holds one header line and one line per pick:
//   fecl-autotune 1
//   <key>\t<words> <vars per wave> <ms (1,1)> <ms (1,2)> <ms (2,1)> <ms (2,2)> <ms (4,1)> <ms (4,2)>
// Keys hold no tab or newline.  Lines that do not parse are skipped and counted.
#pragma once
#include <stdint.h>

#include <string>

#include "ldpc_code.h"

namespace fecl {
namespace tune {

struct Pick {
  int32_t words = 0, vpw = 0;
  double ms[6] = {-1, -1, -1, -1, -1, -1};
};

// FNV-1a over the code's sizes and CSR arrays (the decoder runs on this exact H).
uint64_t code_hash(const Code& c);
// The batch-size bucket of a frame count: up to kFineBucketFrames, the next multiple of 64 (the
// smallest tile), because the layouts' padding to whole tiles decides the pick there (601
// frames fill five 128-frame tiles but three 256-frame ones); above it, the next power of two.
constexpr int64_t kFineBucketFrames = 8192;
int64_t frames_bucket(int64_t frames);
std::string make_key(const std::string& arch, int cus, const Code& c, int algorithm, int msg_format, int64_t frames);

// mode 0: off (lookups miss, stores are dropped); 1: this process; 2: this process and the
// file `path` (read now; each new pick is appended).  Returns false (with *err) when the file
// exists but cannot be read as a cache, or cannot be created.
bool configure(int mode, const char* path, std::string* err);
// Applies FECL_AUTOTUNE_CACHE once, unless configure() ran first.
void configure_from_env();
bool lookup(const std::string& key, Pick* out);
// Keeps the pick (replacing an older one of the same key) and appends it to the file; returns
// false (with *err) when the file cannot be appended (the pick is kept in the process anyway).
bool store(const std::string& key, const Pick& p, std::string* err);
void stats(int64_t* hits, int64_t* misses, int64_t* entries, int64_t* skipped_lines, int64_t* write_failures);
void clear();

}  // namespace tune
}  // namespace fecl
