// Host-side units of the library, built with AddressSanitizer + UndefinedBehaviorSanitizer and
// with ThreadSanitizer (tests/test_sanitizers.py runs both):
//   code construction  every DVB-S2 table, FeCl's generator form of DVB-S2, Gallager and random
//                      matrices, and its encoder (every codeword satisfies H)
//   archive reader     round trips, every truncation, random byte damage
//   autotune cache     lookups and stores from many threads, with and without the file
//   host conversions   FP64 -> FP32 -> FP64 helpers run by concurrent callers sharing threads
//   admission order    which tiles of a serving pool take a round's frames
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <thread>
#include <unistd.h>
#include <vector>

#include "ldpc_admission.h"
#include "ldpc_archive.h"
#include "ldpc_code.h"
#include "ldpc_host_convert.h"
#include "ldpc_tune_cache.h"

using namespace fecl;

static int g_failures = 0;
#define EXPECT(cond, ...)                                            \
  do {                                                               \
    if (!(cond)) {                                                   \
      ++g_failures;                                                  \
      std::printf("FAIL %s:%d: %s | ", __FILE__, __LINE__, #cond);   \
      std::printf(__VA_ARGS__);                                      \
      std::printf("\n");                                             \
    }                                                                \
  } while (0)

static bool satisfies(const Csr& H, const std::vector<uint8_t>& cw) {
  for (int64_t r = 0; r < H.rows(); ++r) {
    int x = 0;
    for (int32_t j = H.ptr[r]; j < H.ptr[r + 1]; ++j) x ^= cw[(size_t)H.idx[j]];
    if (x) return false;
  }
  return true;
}

static void check_generator(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, const char* what,
                            int words) {
  GeneratorForm g;
  const std::string err = generator_form(row_ptr, col, M, N, &g);
  EXPECT(err.empty(), "%s: %s", what, err.c_str());
  if (!err.empty()) return;
  EXPECT(g.M == M && g.N == N && g.K >= N - M && g.H.rows() == M, "%s: sizes", what);
  std::mt19937_64 rng(N * 31 + M);
  std::vector<uint8_t> msg((size_t)g.K), cw((size_t)g.N);
  for (int w = 0; w < words; ++w) {
    for (auto& b : msg) b = (uint8_t)(rng() & 1);
    encode_generator(g, msg.data(), cw.data());
    EXPECT(satisfies(g.H, cw), "%s: codeword %d fails H", what, w);
    EXPECT(std::memcmp(cw.data(), msg.data(), msg.size()) == 0, "%s: not systematic", what);
  }
}

static void test_codes() {
  const int rates[][2] = {{1, 4}, {1, 3}, {2, 5}, {1, 2}, {3, 5}, {2, 3}, {3, 4}, {4, 5}, {5, 6}, {8, 9}, {9, 10}};
  int built = 0;
  for (int n : {64800, 16200}) {
    for (auto& r : rates) {
      int64_t dims[4];
      if (!dvbs2_matrix(n, r[0], r[1], dims, nullptr, nullptr)) continue;
      std::vector<int32_t> rp((size_t)dims[0] + 1), col((size_t)dims[3]);
      EXPECT(dvbs2_matrix(n, r[0], r[1], dims, rp.data(), col.data()), "dvbs2 %d %d/%d", n, r[0], r[1]);
      EXPECT(rp[0] == 0 && rp[(size_t)dims[0]] == dims[3], "dvbs2 %d %d/%d: row pointers", n, r[0], r[1]);
      bool sorted = true;
      for (int64_t i = 0; i < dims[0]; ++i)
        for (int32_t j = rp[i] + 1; j < rp[i + 1]; ++j) sorted &= col[j - 1] < col[j];
      EXPECT(sorted, "dvbs2 %d %d/%d: columns ascending", n, r[0], r[1]);
      ++built;
      if (n == 16200 || (r[0] == 1 && r[1] == 2)) {
        char what[64];
        std::snprintf(what, sizeof(what), "dvbs2 %d %d/%d generator form", n, r[0], r[1]);
        check_generator(rp.data(), col.data(), dims[0], dims[1], what, 2);
      }
    }
  }
  EXPECT(built == 21, "%d DVB-S2 codes built", built);
  Csr gal;
  EXPECT(gallager_matrix(96, 3, 6, 7, &gal).empty(), "gallager");
  check_generator(gal.ptr.data(), gal.idx.data(), gal.rows(), 96, "gallager 96 3 6", 8);
  EXPECT(!gallager_matrix(10, 3, 11, 7, &gal).empty(), "gallager with wr > n is rejected");
  // random sparse matrices, dependent rows included
  std::mt19937_64 rng(99);
  for (int k = 0; k < 40; ++k) {
    const int64_t N = 20 + (int64_t)(rng() % 60), M = 5 + (int64_t)(rng() % (N / 2));
    Csr H;
    for (int64_t r = 0; r < M; ++r) {
      std::vector<int32_t> row;
      for (int32_t c = 0; c < N; ++c)
        if (rng() % 7 == 0) row.push_back(c);
      while (row.size() < 2) {   // every check row has at least two edges
        const int32_t c = (int32_t)(rng() % N);
        if (std::find(row.begin(), row.end(), c) == row.end()) row.insert(std::upper_bound(row.begin(), row.end(), c), c);
      }
      if (k % 5 == 0 && r > 0 && rng() % 3 == 0)   // a copy of the previous row
        row.assign(H.idx.begin() + H.ptr[r - 1], H.idx.end());
      H.idx.insert(H.idx.end(), row.begin(), row.end());
      H.ptr.push_back((int32_t)H.idx.size());
    }
    char what[64];
    std::snprintf(what, sizeof(what), "random %lld x %lld", (long long)M, (long long)N);
    check_generator(H.ptr.data(), H.idx.data(), M, N, what, 4);
  }
}

static void test_archive() {
  Csr gal;
  gallager_matrix(48, 3, 6, 11, &gal);
  GeneratorForm g;
  EXPECT(generator_form(gal.ptr.data(), gal.idx.data(), gal.rows(), 48, &g).empty(), "generator form");
  ArchiveOptions o;
  o.algorithm = 1;
  o.iterations = 7;
  o.rows = 2;
  o.cols = 5;
  o.scaling.assign(10, 0.75);
  std::vector<uint8_t> buf;
  archive_write(g, (int64_t)gal.idx.size(), o, &buf);
  GeneratorForm g2;
  ArchiveOptions o2;
  EXPECT(archive_read(buf.data(), (int64_t)buf.size(), &g2, &o2).empty(), "round trip");
  EXPECT(g2.H.idx == g.H.idx && g2.H.ptr == g.H.ptr && o2.iterations == 7 && o2.scaling == o.scaling,
         "round trip fields");
  int accepted = 0;
  for (size_t len = 0; len < buf.size(); ++len) {   // every truncation is rejected, none crashes
    GeneratorForm gx;
    ArchiveOptions ox;
    accepted += archive_read(buf.data(), (int64_t)len, &gx, &ox).empty() ? 1 : 0;
  }
  EXPECT(accepted == 0, "%d truncated archives accepted", accepted);
  std::mt19937_64 rng(5);
  for (int k = 0; k < 3000; ++k) {   // random damage: parsed or rejected, never out of bounds
    std::vector<uint8_t> d = buf;
    const int flips = 1 + (int)(rng() % 4);
    for (int f = 0; f < flips; ++f) d[rng() % d.size()] ^= (uint8_t)(1u << (rng() % 8));
    GeneratorForm gx;
    ArchiveOptions ox;
    (void)archive_read(d.data(), (int64_t)d.size(), &gx, &ox);
  }
}

static void test_tune_cache() {
  Code c;
  c.N = 64800;
  c.M = 32400;
  c.K = 32400;
  c.E = 226799;
  c.row_ptr = {0, 3, 7};
  c.col = {1, 2, 3, 4, 5, 6, 7};
  char path[] = "/tmp/fecl_tune_cacheXXXXXX";
  const int fd = mkstemp(path);
  EXPECT(fd >= 0, "mkstemp");
  close(fd);
  std::remove(path);
  std::string err;
  EXPECT(tune::configure(2, path, &err), "configure: %s", err.c_str());
  std::vector<std::thread> th;
  std::atomic<int> mismatches{0};
  for (int t = 0; t < 8; ++t) {
    th.emplace_back([&, t] {
      for (int i = 0; i < 200; ++i) {
        const std::string key = tune::make_key("gfx950", 256, c, t % 3, i % 3, 1 + i % 50);
        tune::Pick p;
        p.words = 1 << (i % 3);
        p.vpw = 1 + (t & 1);
        p.ms[0] = i;
        std::string e;
        tune::store(key, p, &e);
        tune::Pick q;
        if (tune::lookup(key, &q) && (q.words < 1 || q.words > 4)) ++mismatches;
      }
    });
  }
  for (auto& x : th) x.join();
  EXPECT(mismatches == 0, "invalid picks read back");
  int64_t hits = 0, misses = 0, entries = 0, skipped = 0, wfail = 0;
  tune::stats(&hits, &misses, &entries, &skipped, &wfail);
  EXPECT(hits == 1600 && entries > 0 && skipped == 0 && wfail == 0, "hits %lld entries %lld skipped %lld",
         (long long)hits, (long long)entries, (long long)skipped);
  // a fresh view of the file holds the same keys
  tune::clear();
  EXPECT(tune::configure(2, path, &err), "reload: %s", err.c_str());
  int64_t entries2 = 0;
  tune::stats(nullptr, nullptr, &entries2, &skipped, nullptr);
  EXPECT(entries2 == entries && skipped == 0, "reloaded %lld of %lld entries", (long long)entries2,
         (long long)entries);
  EXPECT(tune::configure(1, nullptr, &err), "back to process-only");
  std::remove(path);
  // a file that can no longer be appended: the pick stays in the process, the failure counts
  char dir[] = "/tmp/fecl_tune_dirXXXXXX";
  EXPECT(mkdtemp(dir) != nullptr, "mkdtemp");
  const std::string in_dir = std::string(dir) + "/cache.txt";
  EXPECT(tune::configure(2, in_dir.c_str(), &err), "configure in a directory: %s", err.c_str());
  std::remove(in_dir.c_str());
  rmdir(dir);
  tune::Pick p;
  p.words = 2;
  p.vpw = 1;
  const std::string k = tune::make_key("gfx950", 256, c, 0, 0, 64);
  EXPECT(!tune::store(k, p, &err), "append into a removed directory reported");
  tune::Pick q;
  int64_t wf = 0;
  tune::stats(nullptr, nullptr, nullptr, nullptr, &wf);
  EXPECT(tune::lookup(k, &q) && q.words == 2 && wf == 1, "kept in process, failure counted (%lld)", (long long)wf);
  EXPECT(tune::configure(1, nullptr, &err), "back to process-only");
  EXPECT(tune::frames_bucket(1) == 64 && tune::frames_bucket(200) == 256 && tune::frames_bucket(601) == 640 &&
             tune::frames_bucket(1024) == 1024 && tune::frames_bucket(8192) == 8192 &&
             tune::frames_bucket(8193) == 16384 && tune::frames_bucket(20000) == 32768,
         "frame buckets");
}

static void test_host_convert() {
  const size_t n = 3 * 65536 + 123;   // several threads per call
  std::vector<std::thread> callers;
  std::atomic<int> wrong{0};
  for (int t = 0; t < 6; ++t) {
    callers.emplace_back([&, t] {
      ++engine::g_host_calls;
      std::mt19937_64 rng(t);
      std::normal_distribution<double> nd(0.0, 10.0);
      std::vector<double> src(n), back(n), syst(n / 2);
      std::vector<float> mid(n);
      for (auto& x : src) x = nd(rng);
      for (auto& x : syst) x = nd(rng) * 1e-3;
      for (int rep = 0; rep < 3; ++rep) {
        engine::to_float(src.data(), mid.data(), n);
        engine::to_double(mid.data(), back.data(), n);
        for (size_t i = 0; i < n; ++i)
          if (back[i] != (double)(float)src[i]) {
            ++wrong;
            break;
          }
        // two frames of N = n / 2 with K = n / 4 a-priori values each
        const int64_t N = (int64_t)(n / 2), K = N / 2;
        engine::to_float_plus(src.data(), syst.data(), mid.data(), 2, N, K);
        for (int64_t f = 0; f < 2; ++f)
          for (int64_t v = 0; v < N; v += 997) {
            const double l = src[(size_t)(f * N + v)];
            const float want = v < K ? (float)((double)(float)l + syst[(size_t)(f * K + v)]) : (float)l;
            if (mid[(size_t)(f * N + v)] != want) ++wrong;
          }
        std::vector<uint8_t> a(n), b(n);
        for (size_t i = 0; i < n; ++i) a[i] = (uint8_t)(i * 7 + t);
        engine::copy_bytes(a.data(), b.data(), n);
        if (a != b) ++wrong;
      }
      --engine::g_host_calls;
    });
  }
  for (auto& x : callers) x.join();
  EXPECT(wrong == 0, "%d conversion mismatches", wrong.load());
}

static std::vector<int> order_of(const std::vector<int64_t>& free_slots, const std::vector<int64_t>& used,
                                 int64_t want, int max_occupied) {
  std::vector<int> o;
  admission_order(free_slots, used, want, max_occupied, &o);
  return o;
}

static void test_admission_order() {
  using V = std::vector<int>;
  // a nearly full tile (4 free), a sparse one (54 free), two empty ones
  const std::vector<int64_t> f1 = {4, 54, 64, 64}, u1 = {60, 10, 0, 0};
  EXPECT(order_of(f1, u1, 3, 2) == V({0, 1, 2, 3}), "a small round goes to the fullest tile that takes it");
  EXPECT(order_of(f1, u1, 4, 2) == V({0, 1, 2, 3}), "exactly the free slots of the fullest tile");
  EXPECT(order_of(f1, u1, 10, 2) == V({1, 0, 2, 3}), "only the sparse tile takes 10 frames");
  EXPECT(order_of(f1, u1, 128, 2) == V({1, 0, 2, 3}), "no tile takes the round: most room first");
  // at most max_occupied tiles that hold frames, the fullest fitting one first
  const std::vector<int64_t> f2 = {10, 20, 30, 5, 64}, u2 = {54, 44, 34, 59, 0};
  EXPECT(order_of(f2, u2, 25, 2) == V({2, 1, 4}), "25 frames: only tile 2 has room");
  EXPECT(order_of(f2, u2, 8, 2) == V({0, 2, 4}), "8 frames: tile 0 is the fullest with room");
  EXPECT(order_of(f2, u2, 8, 4) == V({0, 2, 1, 3, 4}), "8 frames, four occupied tiles allowed");
  // full tiles take nothing; a pool with no frame fills empty tiles in index order
  EXPECT(order_of({0, 5}, {64, 59}, 1, 2) == V({1}), "full tile skipped");
  EXPECT(order_of({64, 64}, {0, 0}, 200, 2) == V({0, 1}), "empty pool");
  // equal free counts keep index order among themselves
  EXPECT(order_of({8, 8, 3}, {56, 56, 61}, 5, 2) == V({1, 0}), "ties: the later of the tiles with room");
  // random pools: at most max_occupied occupied tiles, then every empty tile in index order; the
  // first occupied tile has room for the round if any tile has
  std::mt19937_64 rng(7);
  for (int iter = 0; iter < 2000; ++iter) {
    const int T = 1 + (int)(rng() % 12), tw = 64;
    std::vector<int64_t> fr(T), us(T);
    for (int t = 0; t < T; ++t) {
      us[t] = (int64_t)(rng() % 3 == 0 ? 0 : rng() % (tw + 1));
      fr[t] = tw - us[t];
    }
    const int64_t want = 1 + (int64_t)(rng() % 100);
    const int cap = 1 + (int)(rng() % 3);
    const std::vector<int> o = order_of(fr, us, want, cap);
    int occupied = 0, empties = 0;
    bool ok = true;
    for (size_t i = 0; i < o.size(); ++i) {
      if (us[o[i]] > 0) {
        ok &= fr[o[i]] > 0 && empties == 0;
        ++occupied;
      } else {
        ok &= (i == 0 || us[o[i - 1]] > 0 || o[i - 1] < o[i]);
        ++empties;
      }
    }
    int64_t fits = -1, most = -1;
    for (int t = 0; t < T; ++t) {
      if (us[t] == 0 || fr[t] == 0) continue;
      if (fr[t] >= want && (fits < 0 || fr[t] < fits)) fits = fr[t];
      most = std::max(most, fr[t]);
    }
    const int total_empty = (int)std::count(us.begin(), us.end(), (int64_t)0);
    ok &= occupied <= cap && empties == total_empty;
    if (occupied > 0) ok &= fr[o[0]] == (fits >= 0 ? fits : most);
    EXPECT(ok, "random pool %d (T=%d want=%lld cap=%d)", iter, T, (long long)want, cap);
    if (!ok) break;
  }
}

int main() {
  test_codes();
  test_archive();
  test_tune_cache();
  test_host_convert();
  test_admission_order();
  std::printf("%d failures\n", g_failures);
  return g_failures ? 1 : 0;
}
