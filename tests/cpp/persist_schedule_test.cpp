// Checks the persistent decode's work queue (fecl_amd/csrc/ldpc_persist.h) on the CPU: walked
// in queue order by a single worker (every item finishing before the next starts), each item's
// wait must already be satisfied; so with any number of workgroups taking items in queue
// order, every item waits only for items taken before it, and the decode cannot deadlock.
// The model follows the kernel's bookkeeping: per-tile check / variable / syndrome counters,
// tile controls (finished by the last syndrome item of a tile), checkpoint plans, moves,
// results and tile-flag items, with tiles retiring, codes finishing and compactions going
// ahead at random.  Also checks that the queue holds each (kind, code, tile, chunk, step)
// exactly once.  Built with hipcc (host code only; no GPU needed):  make -C tests/cpp
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <tuple>
#include <vector>

#include "../../fecl_amd/csrc/ldpc_persist.h"

using namespace fecl::persist;

static int failures = 0;
#define CHECK(c, ...)                      \
  do {                                     \
    if (!(c)) {                            \
      std::printf("FAIL: " __VA_ARGS__);   \
      std::printf("\n");                   \
      ++failures;                          \
    }                                      \
  } while (0)

struct CodeState {
  CodeShape sh;
  std::vector<int> cn, vn, syn, ctrl, retired;
  int ctrl_all = 0, plan_done = 0, res_done = 0, mv_done = 0, retile_pre = 0, retile_done = 0;
  int go = 0, new_tiles = 0, cur_tiles = 0, live_tiles = 0, done = 0;
};

static void run_case(const std::vector<CodeShape>& shapes, uint32_t seed) {
  Queue q;
  build_queue(shapes.data(), (int)shapes.size(), &q);
  std::mt19937 rng(seed);
  std::vector<CodeState> cs(shapes.size());
  for (size_t c = 0; c < shapes.size(); ++c) {
    CodeState& s = cs[c];
    s.sh = shapes[c];
    const int T = s.sh.tiles;
    s.cn.assign(T, 0); s.vn.assign(T, 0); s.syn.assign(T, 0); s.ctrl.assign(T, 0); s.retired.assign(T, 0);
    s.cur_tiles = T;
    s.live_tiles = T;
  }
  // every (kind, code, tile, chunk, step) once
  std::set<std::tuple<int, int, int, int, int>> seen;
  for (int64_t i = 0; i < q.total; ++i) {
    const Item it = decode_item(q, i);
    CHECK(it.kind >= 0, "item %lld does not decode", (long long)i);
    if (it.kind < 0) return;
    CHECK(seen.insert(std::make_tuple(it.kind, it.code, it.tile, it.chunk, it.step)).second,
          "item %lld repeats", (long long)i);
    CodeState& s = cs[it.code];
    const int T = s.sh.tiles, t = it.tile, st = it.step;
    CHECK(it.code < (int)shapes.size() && t < T && st < q.imax, "item %lld out of range", (long long)i);
    if (s.done) continue;   // the kernel skips a finished code's items without counting them
    bool ok = false;
    const int cp = cp_index(st);
    switch (it.kind) {
      case K_CN:
        ok = st == 0 || (is_checkpoint(q, st - 1) ? s.retile_done >= cp_index(st - 1) + 1 : s.ctrl[t] >= st);
        break;
      case K_VN: ok = s.cn[t] >= (st + 1) * s.sh.cn_chunks; break;
      case K_SYN: ok = s.vn[t] >= (st + 1) * s.sh.vn_chunks; break;
      case K_PLAN: ok = s.ctrl_all >= (st + 1) * T; break;
      case K_MOVE:
      case K_RES: ok = s.plan_done >= cp + 1; break;
      case K_RETILE:
        ok = s.mv_done >= (cp + 1) * T * s.sh.mv_chunks && s.res_done >= (cp + 1) * T * s.sh.res_chunks;
        break;
    }
    CHECK(ok, "seed %u: item %lld (kind %d code %d tile %d chunk %d step %d) would wait for a later item", seed,
          (long long)i, it.kind, it.code, t, it.chunk, st);
    switch (it.kind) {
      case K_CN: ++s.cn[t]; break;
      case K_VN: ++s.vn[t]; break;
      case K_SYN:
        if (++s.syn[t] == (st + 1) * s.sh.syn_chunks) {
          // tile control: a live tile may retire (frames finish at random; none after the
          // code's last step)
          if (t < s.cur_tiles && !s.retired[t] && st < s.sh.iters) {
            const bool dies = st + 1 >= s.sh.iters || rng() % 5 == 0;
            if (dies) {
              s.retired[t] = 1;
              if (--s.live_tiles == 0) s.done = 1;
            }
          }
          ++s.ctrl[t];
          ++s.ctrl_all;
        }
        break;
      case K_PLAN: {
        const int n = s.cur_tiles;
        s.go = (st + kMinSteps < s.sh.iters) && n > 1 && rng() % 2 == 0;
        s.new_tiles = s.go ? 1 + (int)(rng() % (3 * n / 4 > 0 ? 3 * n / 4 : 1)) : n;
        ++s.plan_done;
        break;
      }
      case K_MOVE: ++s.mv_done; break;
      case K_RES: ++s.res_done; break;
      case K_RETILE:
        if (s.go) s.retired[t] = t < s.new_tiles ? 0 : 1;
        if (++s.retile_pre == (cp + 1) * T) {
          if (s.go) {
            s.cur_tiles = s.new_tiles;
            s.live_tiles = s.new_tiles;
          }
          s.retile_done = cp + 1;
        }
        break;
    }
  }
  for (size_t c = 0; c < shapes.size(); ++c) CHECK(cs[c].done, "seed %u: code %zu never finished", seed, c);
}

int main() {
  std::mt19937 rng(7);
  int cases = 0;
  for (int trial = 0; trial < 300; ++trial) {
    const int n = 1 + (int)(rng() % 7);
    std::vector<CodeShape> shapes(n);
    for (auto& sh : shapes) {
      sh.tiles = 1 + (int)(rng() % 6);
      sh.iters = 1 + (int)(rng() % 30);
      sh.cn_chunks = 1 + (int)(rng() % 5);
      sh.vn_chunks = 1 + (int)(rng() % 5);
      sh.syn_chunks = 1 + (int)(rng() % 3);
      sh.mv_chunks = 1 + (int)(rng() % 4);
      sh.res_chunks = 1 + (int)(rng() % 3);
    }
    run_case(shapes, (uint32_t)trial);
    ++cases;
  }
  // the seven DVB-S2 rates of the uneven workload, 50 iterations
  const int64_t dims[7][4] = {{48600, 64800, 16200, 194399}, {32400, 64800, 32400, 226799},
                              {25920, 64800, 38880, 285119}, {21600, 64800, 43200, 215999},
                              {16200, 64800, 48600, 226799}, {10800, 64800, 54000, 237599},
                              {6480, 64800, 58320, 194399}};
  std::vector<CodeShape> dvb(7);
  for (int c = 0; c < 7; ++c) {
    chunk_counts(dims[c][0], dims[c][1], dims[c][2], dims[c][3], &dvb[c]);
    dvb[c].tiles = 1 + c % 4;
    dvb[c].iters = 50;
  }
  for (uint32_t seed = 0; seed < 5; ++seed) run_case(dvb, 1000 + seed);
  std::printf("%d random cases, 5 DVB-S2 cases, %d failures\n", cases, failures);
  return failures ? 1 : 0;
}
