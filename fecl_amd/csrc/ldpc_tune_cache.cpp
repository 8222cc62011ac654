// Autotune cache (see ldpc_tune_cache.h).
#include "ldpc_tune_cache.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <unordered_map>

namespace fecl {
namespace tune {

namespace {

const char kHeader[] = "fecl-autotune 1";

struct Cache {
  std::mutex mu;
  int mode = 1;
  bool configured = false;   // configure() ran (the environment is then ignored)
  std::string path;
  std::unordered_map<std::string, Pick> map;
  int64_t hits = 0, misses = 0, skipped = 0, write_failures = 0;
};

Cache& cache() {
  static Cache c;
  return c;
}

uint64_t fnv(uint64_t h, const void* p, size_t n) {
  const unsigned char* b = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= b[i];
    h *= 1099511628211ull;
  }
  return h;
}

std::string format_line(const std::string& key, const Pick& p) {
  std::ostringstream os;
  os.precision(6);
  os << key << '\t' << p.words << ' ' << p.vpw;
  for (double m : p.ms) os << ' ' << m;
  os << '\n';
  return os.str();
}

bool valid(const Pick& p) { return (p.words == 1 || p.words == 2 || p.words == 4) && (p.vpw == 1 || p.vpw == 2); }

// Reads `path` into c.map (caller holds c.mu).  A missing file is an empty cache.
bool load_locked(Cache& c, const std::string& path, std::string* err) {
  std::ifstream in(path);
  if (!in) return true;
  std::string line;
  if (!std::getline(in, line)) return true;   // empty file
  if (line != kHeader) {
    if (err) *err = "autotune cache " + path + ": not a cache file (first line is not '" + kHeader + "')";
    return false;
  }
  while (std::getline(in, line)) {
    const size_t tab = line.find('\t');
    if (tab == std::string::npos || tab == 0) {
      if (!line.empty()) ++c.skipped;
      continue;
    }
    std::istringstream vs(line.substr(tab + 1));
    Pick p;
    vs >> p.words >> p.vpw;
    for (double& m : p.ms) vs >> m;
    if (!vs || !valid(p)) {
      ++c.skipped;
      continue;
    }
    c.map[line.substr(0, tab)] = p;
  }
  return true;
}

}  // namespace

uint64_t code_hash(const Code& c) {
  uint64_t h = 1469598103934665603ull;
  const int64_t dims[4] = {c.M, c.N, c.K, c.E};
  h = fnv(h, dims, sizeof(dims));
  h = fnv(h, c.row_ptr.data(), c.row_ptr.size() * sizeof(int32_t));
  h = fnv(h, c.col.data(), c.col.size() * sizeof(int32_t));
  return h;
}

int64_t frames_bucket(int64_t frames) {
  if (frames <= kFineBucketFrames) return std::max<int64_t>(1, (frames + 63) / 64) * 64;
  int64_t b = kFineBucketFrames;
  while (b < frames && b < (int64_t(1) << 40)) b <<= 1;
  return b;
}

std::string make_key(const std::string& arch, int cus, const Code& c, int algorithm, int msg_format, int64_t frames) {
  std::string a;
  for (char ch : arch) a += (ch == '\t' || ch == '\n' || ch == ' ') ? '_' : ch;
  char buf[256];
  std::snprintf(buf, sizeof(buf), "/%d|n%lld m%lld e%lld h%016llx|alg%d|fmt%d|f%lld", cus, (long long)c.N,
                (long long)c.M, (long long)c.E, (unsigned long long)code_hash(c), algorithm, msg_format,
                (long long)frames_bucket(frames));
  return a + buf;
}

bool configure(int mode, const char* path, std::string* err) {
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  if (mode < 0 || mode > 2) {
    if (err) *err = "autotune cache mode must be 0 (off), 1 (process) or 2 (process and file)";
    return false;
  }
  if (mode == 2 && (!path || !*path)) {
    if (err) *err = "autotune cache mode 2 needs a file path";
    return false;
  }
  c.configured = true;
  c.mode = mode;
  c.path = mode == 2 ? path : "";
  if (mode == 2) {
    if (!load_locked(c, c.path, err)) {
      c.mode = 1;
      c.path.clear();
      return false;
    }
    // a missing or empty file gets its header now, so that later appends need no check
    std::ifstream probe(c.path, std::ios::ate);
    if (!probe || probe.tellg() == 0) {
      std::ofstream out(c.path, std::ios::trunc);
      if (!out || !(out << kHeader << '\n')) {
        if (err) *err = "autotune cache " + c.path + ": cannot create the file";
        c.mode = 1;
        c.path.clear();
        return false;
      }
    }
  }
  return true;
}

void configure_from_env() {
  Cache& c = cache();
  {
    std::lock_guard<std::mutex> lk(c.mu);
    if (c.configured) return;
    c.configured = true;
  }
  const char* p = std::getenv("FECL_AUTOTUNE_CACHE");
  if (p && *p) {
    std::string err;
    if (!configure(2, p, &err)) std::fprintf(stderr, "fecl: %s (kept in process only)\n", err.c_str());
  }
}

bool lookup(const std::string& key, Pick* out) {
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  if (c.mode == 0) return false;
  auto it = c.map.find(key);
  if (it == c.map.end()) {
    ++c.misses;
    return false;
  }
  ++c.hits;
  *out = it->second;
  return true;
}

bool store(const std::string& key, const Pick& p, std::string* err) {
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  if (c.mode == 0 || !valid(p) || key.find('\t') != std::string::npos || key.find('\n') != std::string::npos)
    return true;
  c.map[key] = p;
  if (c.mode == 2) {
    std::ofstream out(c.path, std::ios::app);
    if (!out || !(out << format_line(key, p))) {
      ++c.write_failures;
      if (err) *err = "autotune cache " + c.path + ": cannot append";
      return false;
    }
  }
  return true;
}

void stats(int64_t* hits, int64_t* misses, int64_t* entries, int64_t* skipped_lines, int64_t* write_failures) {
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  if (hits) *hits = c.hits;
  if (misses) *misses = c.misses;
  if (entries) *entries = (int64_t)c.map.size();
  if (skipped_lines) *skipped_lines = c.skipped;
  if (write_failures) *write_failures = c.write_failures;
}

void clear() {
  Cache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  c.map.clear();
  c.hits = c.misses = c.skipped = c.write_failures = 0;
}

}  // namespace tune
}  // namespace fecl
