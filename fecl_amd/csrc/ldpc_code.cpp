#include "ldpc_code.h"

#include <algorithm>
#include <numeric>
#include <random>

#include "dvbs2_tables.h"

namespace fecl {

std::string build_code(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, Code* out,
                       int64_t k) {
  if (!row_ptr || !col) return "null parity-check arrays";
  if (M < 1 || N < 1 || M >= N) return "parity-check matrix must have 1 <= rows < cols";
  if (row_ptr[0] != 0) return "row_ptr[0] must be 0";
  if (k >= 0 && (k < N - M || k >= N)) return "message size out of range";
  Code c;
  c.M = M;
  c.N = N;
  c.K = k >= 0 ? k : N - M;
  c.row_ptr.assign(row_ptr, row_ptr + M + 1);
  for (int64_t r = 0; r < M; ++r) {
    int32_t d = row_ptr[r + 1] - row_ptr[r];
    if (d < 2) return "every check row needs at least two edges";
    c.max_row_degree = std::max(c.max_row_degree, d);
  }
  c.E = row_ptr[M];
  c.col.assign(col, col + c.E);
  std::vector<int32_t> deg(N, 0);
  for (int64_t e = 0; e < c.E; ++e) {
    if (col[e] < 0 || col[e] >= N) return "column index out of range";
    deg[col[e]]++;
  }
  c.var_ptr.assign(N + 1, 0);
  for (int64_t v = 0; v < N; ++v) {
    c.var_ptr[v + 1] = c.var_ptr[v] + deg[v];
    c.max_var_degree = std::max(c.max_var_degree, deg[v]);
  }
  c.var_edges.assign(c.E, 0);
  std::vector<int32_t> fill(c.var_ptr.begin(), c.var_ptr.end() - 1);
  for (int64_t e = 0; e < c.E; ++e) c.var_edges[fill[col[e]]++] = (int32_t)e;  // ascending e
  // Staircase test: row r holds exactly the parity columns K+r-1 (r > 0) and K+r.
  bool st = true;
  for (int64_t r = 0; r < M && st; ++r) {
    int n_par = 0;
    for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
      int64_t v = col[e];
      if (v >= c.K) {
        ++n_par;
        if (v != c.K + r && v != c.K + r - 1) st = false;
      }
    }
    if (n_par != (r == 0 ? 1 : 2)) st = false;
  }
  c.staircase = st;
  *out = std::move(c);
  return "";
}

bool dvbs2_matrix(int n, int rate_num, int rate_den, int64_t* dims, int32_t* row_ptr, int32_t* col) {
  const fecl_dvbs2_code_entry* ce = nullptr;
  for (const auto& e : fecl_dvbs2_codes)
    if (e.n == n && e.rate_num == rate_num && e.rate_den == rate_den) ce = &e;
  if (!ce) return false;
  const int64_t q = ce->q, M = 360 * q, K = n - M;
  // Information bit i of group g = i/360 joins check (a + (i mod 360) q) mod M for every
  // address a of table row g; parity bit K+j joins checks j and j+1.  Rows list columns in
  // ascending order, which is the order FeCl's transpose() produces.
  std::vector<int32_t> deg(M, 0);
  struct Group { const uint16_t* addr; int count; };
  std::vector<Group> groups;
  const uint16_t* rec = fecl_dvbs2_addr + ce->offset;
  for (int g = 0; g < ce->table_rows; ++g) {
    groups.push_back({rec + 1, rec[0]});
    rec += 1 + rec[0];
  }
  for (const auto& g : groups)
    for (int k = 0; k < 360; ++k)
      for (int j = 0; j < g.count; ++j) deg[(g.addr[j] + k * q) % M]++;
  for (int64_t j = 0; j < M; ++j) {
    deg[j]++;
    if (j + 1 < M) deg[j + 1]++;
  }
  int64_t E = 0;
  for (auto d : deg) E += d;
  if (dims) {
    dims[0] = M; dims[1] = n; dims[2] = K; dims[3] = E;
  }
  if (!row_ptr || !col) return true;
  std::vector<int64_t> pos(M);
  row_ptr[0] = 0;
  for (int64_t r = 0; r < M; ++r) {
    row_ptr[r + 1] = row_ptr[r] + deg[r];
    pos[r] = row_ptr[r];
  }
  for (size_t g = 0; g < groups.size(); ++g)
    for (int k = 0; k < 360; ++k) {
      const int32_t v = (int32_t)(g * 360 + k);
      for (int j = 0; j < groups[g].count; ++j) col[pos[(groups[g].addr[j] + k * q) % M]++] = v;
    }
  for (int64_t j = 0; j < M; ++j) {
    col[pos[j]++] = (int32_t)(K + j);
    if (j + 1 < M) col[pos[j + 1]++] = (int32_t)(K + j);
  }
  return true;
}

std::string scaling_map_to_table(const std::unordered_map<size_t, std::vector<double>>& map,
                                 const Code& c, int64_t iterations,
                                 std::vector<double>* table, int32_t* rows, int32_t* cols) {
  const size_t maxDegree = (size_t)c.max_row_degree;
  std::vector<std::vector<double>> vec;
  auto def = map.find(0);
  if (def != map.end()) {
    if (def->second.size() != 1 && (int64_t)def->second.size() != iterations)
      return "Wrong size for scaling factor";
    vec.resize(def->second.size());
    for (size_t i = 0; i < vec.size(); ++i) vec[i].assign(maxDegree - 1, def->second[i]);
    for (const auto& kv : map) {
      if (kv.first >= 2 && kv.first - 2 < maxDegree - 1) {
        if (kv.second.size() != vec.size()) return "Wrong size for scaling factor";
        for (size_t j = 0; j < vec.size(); ++j) vec[j][kv.first - 2] = kv.second[j];
      }
    }
  } else {
    size_t length = 0;
    for (const auto& kv : map) length = std::max(kv.second.size(), length);
    vec.assign(length, std::vector<double>(maxDegree - 1, 0.0));
    if (length != 1 && (int64_t)length != iterations) return "Wrong size for scaling factor";
    for (int64_t r = 0; r < c.M; ++r) {
      size_t d = (size_t)(c.row_ptr[r + 1] - c.row_ptr[r]);
      auto it = map.find(d);
      if (it == map.end()) return "Scaling factor not defined and no default";
      if (it->second.size() != length) return "Wrong size for scaling factor";
      for (size_t j = 0; j < vec.size(); ++j) vec[j][d - 2] = it->second[j];
    }
  }
  *rows = (int32_t)vec.size();
  *cols = (int32_t)(maxDegree - 1);
  table->clear();
  for (auto& r : vec) table->insert(table->end(), r.begin(), r.end());
  return "";
}

// ---------------------------------------------------------------------------------------------
// Generator form (Structure::computeGeneratorMatrix, src/detail/Ldpc.cpp:216-331)
//
// FeCl works on SparseBitMatrix/BitMatrix objects and physically swaps columns in every row.
// Here every row keeps its original column ids; columns carry a label permutation
// (lab2orig/orig2lab) and rows a position permutation (pos2row/row2pos), which is all a column
// swap, a column move or a row swap changes.  The greedy choice of the next triangular column
// scans all columns in FeCl (O(N) per step, O(N*M) for the row partition); here it takes the
// largest label from per-count buckets, and the partition is replayed from the column's
// incidence list, so a 64800-column code is transformed in milliseconds with the same result.
namespace {

using Bits = std::vector<uint64_t>;
inline bool bit(const Bits& b, int64_t i) { return (b[(size_t)(i >> 6)] >> (i & 63)) & 1; }
inline void flip(Bits& b, int64_t i) { b[(size_t)(i >> 6)] ^= uint64_t(1) << (i & 63); }

void append_row(Csr* m, std::vector<int32_t>& cols) {
  std::sort(cols.begin(), cols.end());
  m->idx.insert(m->idx.end(), cols.begin(), cols.end());
  m->ptr.push_back((int32_t)m->idx.size());
  cols.clear();
}

}  // namespace

std::string generator_form(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N,
                           GeneratorForm* out) {
  if (!row_ptr || !col) return "null parity-check arrays";
  if (M < 1 || N < 1 || M >= N) return "parity-check matrix must have 1 <= rows < cols";
  if (row_ptr[0] != 0) return "row_ptr[0] must be 0";
  for (int64_t r = 0; r < M; ++r) {
    if (row_ptr[r + 1] - row_ptr[r] < 2) return "every check row needs at least two edges";
    for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
      if (col[e] < 0 || col[e] >= N) return "column index out of range";
      if (e > row_ptr[r] && col[e] <= col[e - 1]) return "columns must be ascending within each row";
    }
  }
  const int64_t E = row_ptr[M];
  // rows holding each original column, ascending
  std::vector<int32_t> cptr((size_t)N + 1, 0), crow((size_t)E);
  for (int64_t e = 0; e < E; ++e) cptr[(size_t)col[e] + 1]++;
  for (int64_t v = 0; v < N; ++v) cptr[v + 1] += cptr[v];
  {
    std::vector<int32_t> fill(cptr.begin(), cptr.end() - 1);
    for (int64_t r = 0; r < M; ++r)
      for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) crow[(size_t)fill[col[e]]++] = (int32_t)r;
  }
  std::vector<int32_t> pos2row((size_t)M), row2pos((size_t)M), lab2orig((size_t)N), orig2lab((size_t)N);
  std::iota(pos2row.begin(), pos2row.end(), 0);
  std::iota(row2pos.begin(), row2pos.end(), 0);
  std::iota(lab2orig.begin(), lab2orig.end(), 0);
  std::iota(orig2lab.begin(), orig2lab.end(), 0);
  auto has = [&](int32_t r, int32_t o) { return std::binary_search(col + row_ptr[r], col + row_ptr[r + 1], o); };
  auto swap_pos = [&](int64_t p, int64_t q) {
    std::swap(pos2row[p], pos2row[q]);
    row2pos[pos2row[p]] = (int32_t)p;
    row2pos[pos2row[q]] = (int32_t)q;
  };
  auto swap_lab = [&](int64_t x, int64_t y) {
    std::swap(lab2orig[x], lab2orig[y]);
    orig2lab[lab2orig[x]] = (int32_t)x;
    orig2lab[lab2orig[y]] = (int32_t)y;
  };

  // 1. Triangular part, chosen from the last column down (src/detail/Ldpc.cpp:218-256).  cnt[o]
  // counts the rows at positions [0, maxRow) holding original column o.  The next column is the
  // largest unprocessed label of count 1, else the largest of the smallest count >= 1; it moves
  // to label i-1 and its rows to the bottom of [0, maxRow), in the order FeCl's in-place
  // partition leaves them.
  std::vector<int32_t> cnt((size_t)N);
  int32_t maxc = 0;
  for (int64_t o = 0; o < N; ++o) {
    cnt[o] = cptr[o + 1] - cptr[o];
    maxc = std::max(maxc, cnt[o]);
  }
  // Lazy max-heaps of labels per count: an entry (c, l) is live while l is unprocessed and the
  // column at label l still has count c; stale entries are dropped when they reach the top.
  std::vector<std::vector<int32_t>> heap((size_t)maxc + 1);
  for (int64_t l = 0; l < N; ++l)
    if (cnt[l] > 0) heap[cnt[l]].push_back((int32_t)l);
  for (auto& hp : heap) std::make_heap(hp.begin(), hp.end());
  auto push = [&](int32_t c, int32_t l) {
    heap[c].push_back(l);
    std::push_heap(heap[c].begin(), heap[c].end());
  };
  int64_t maxRow = M, tSize = 0;
  std::vector<int32_t> marked;
  for (int64_t i = N; i > 0; --i) {
    const int32_t top = (int32_t)(i - 1);
    int32_t minIdx = -1, minValue = 0;
    for (int32_t c = 1; c <= maxc && minIdx < 0; ++c) {
      auto& hp = heap[c];
      while (!hp.empty() && (hp.front() > top || cnt[lab2orig[hp.front()]] != c)) {
        std::pop_heap(hp.begin(), hp.end());
        hp.pop_back();
      }
      if (!hp.empty()) {
        minIdx = hp.front();
        minValue = c;
      }
    }
    if (minIdx < 0) return "check matrix rows are empty after triangularisation";
    if (minIdx != top) {
      swap_lab(minIdx, top);
      const int32_t cm = cnt[lab2orig[minIdx]];
      if (cm > 0) push(cm, minIdx);
    }
    const int32_t o = lab2orig[top];
    marked.clear();
    for (int32_t k = cptr[o]; k < cptr[o + 1]; ++k)
      if (row2pos[crow[k]] < maxRow) marked.push_back(row2pos[crow[k]]);
    std::sort(marked.begin(), marked.end());
    // FeCl scans positions upward; a row holding the column is swapped with the last row of
    // the shrinking range and the position is examined again.
    int64_t j = maxRow, cursor = 0;
    for (;;) {
      auto it = std::lower_bound(marked.begin(), marked.end(), (int32_t)cursor);
      if (it == marked.end() || *it >= j) break;
      const int32_t p = *it;
      marked.erase(it);
      --j;
      auto jt = std::lower_bound(marked.begin(), marked.end(), (int32_t)j);
      const bool back = jt != marked.end() && *jt == j;
      if (back) marked.erase(jt);
      swap_pos(p, j);
      if (back) marked.insert(std::lower_bound(marked.begin(), marked.end(), p), p);
      cursor = p;
    }
    if (maxRow - j != minValue) return "internal error: triangular column count";
    maxRow = j;
    if (maxRow == 0) {
      tSize = N - i + 1;
      break;
    }
    for (int64_t p = maxRow; p < maxRow + minValue; ++p) {
      const int32_t r = pos2row[p];
      for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) {
        const int32_t c = col[e], l = orig2lab[c];
        if (--cnt[c] > 0 && l < top) push(cnt[c], l);
      }
    }
  }

  // 2. Put column N-tSize+i on the diagonal of row i, from the first row below holding it
  // (src/detail/Ldpc.cpp:258-269).
  for (int64_t i = 0; i < tSize; ++i) {
    const int32_t o = lab2orig[i + N - tSize];
    if (has(pos2row[i], o)) continue;
    int64_t best = M;
    for (int32_t k = cptr[o]; k < cptr[o + 1]; ++k) {
      const int64_t p = row2pos[crow[k]];
      if (p > i && p < best) best = p;
    }
    if (best < M) swap_pos(i, best);
  }

  // 3. The rows below the triangle, dense, with the triangular columns eliminated
  // (src/detail/Ldpc.cpp:271-279).  Bits are indexed by original column.
  const int64_t g = M - tSize, W = (N + 63) / 64;
  std::vector<Bits> cde((size_t)g, Bits((size_t)W, 0));
  for (int64_t k = 0; k < g; ++k) {
    const int32_t r = pos2row[tSize + k];
    for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) flip(cde[k], col[e]);
  }
  for (int64_t i = tSize - 1; i >= 0; --i) {
    const int32_t o = lab2orig[i + N - tSize], r = pos2row[i];
    for (int64_t k = 0; k < g; ++k)
      if (bit(cde[k], o))
        for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) flip(cde[k], col[e]);
  }
  auto xor_row = [&](Bits& dst, const Bits& src) {
    for (int64_t w = 0; w < W; ++w) dst[w] ^= src[w];
  };
  auto first_label = [&](const Bits& b) {
    int64_t best = -1;
    for (int64_t w = 0; w < W; ++w)
      for (uint64_t x = b[w]; x; x &= x - 1) {
        const int64_t l = orig2lab[w * 64 + __builtin_ctzll(x)];
        if (best < 0 || l < best) best = l;
      }
    return best;
  };

  // 4. Forward elimination over the gap columns K .. N-tSize-1; a dependent row adds a message
  // column (src/detail/Ldpc.cpp:281-314).
  int64_t K = N - M;
  for (int64_t i = 0; i < N - tSize - K; ++i) {
    const int64_t lab = i + K;
    bool found = false;
    for (int64_t k = i; k < g; ++k)
      if (bit(cde[k], lab2orig[lab])) {
        std::swap(cde[k], cde[i]);
        found = true;
        break;
      }
    if (!found)
      for (int64_t k = i; k < g; ++k) {
        const int64_t f = first_label(cde[k]);
        if (f >= 0) {
          swap_lab(f, lab);
          std::swap(cde[k], cde[i]);
          found = true;
          break;
        }
      }
    if (!found) {
      // every row from i on is zero: column lab becomes message column K
      const int32_t o = lab2orig[lab];
      for (int64_t l = lab; l > K; --l) {
        lab2orig[l] = lab2orig[l - 1];
        orig2lab[lab2orig[l]] = (int32_t)l;
      }
      lab2orig[K] = o;
      orig2lab[o] = (int32_t)K;
      ++K;
      --i;
      continue;
    }
    const int32_t o = lab2orig[lab];
    for (int64_t k = i + 1; k < g; ++k)
      if (bit(cde[k], o)) xor_row(cde[k], cde[i]);
  }
  const int64_t P = N - K - tSize;
  // 5. Back substitution (src/detail/Ldpc.cpp:316-322).  Rows P .. g-1 are zero by step 4.
  for (int64_t i = P - 1; i >= 0; --i) {
    const int32_t o = lab2orig[i + K];
    for (int64_t k = i - 1; k >= 0; --k)
      if (bit(cde[k], o)) xor_row(cde[k], cde[i]);
  }

  // 6. H_ and the encoder blocks (src/detail/Ldpc.cpp:324-330).
  GeneratorForm f;
  f.M = M;
  f.N = N;
  f.K = K;
  f.t = tSize;
  f.P = P;
  f.H.idx.reserve((size_t)E);
  std::vector<int32_t> tmp, ta, tb;
  std::vector<int32_t> tcount((size_t)tSize + 1, 0);
  for (int64_t p = 0; p < M; ++p) {
    const int32_t r = pos2row[p];
    for (int32_t e = row_ptr[r]; e < row_ptr[r + 1]; ++e) tmp.push_back(orig2lab[col[e]]);
    if (p < tSize) {
      for (int32_t l : tmp) {
        if (l < K) ta.push_back(l);
        else if (l < K + P) tb.push_back((int32_t)(l - K));
        else if (l >= N - tSize) tcount[l - (N - tSize) + 1]++;
      }
      append_row(&f.a, ta);
      append_row(&f.b, tb);
    }
    append_row(&f.H, tmp);
  }
  f.tt.ptr.assign((size_t)tSize + 1, 0);
  for (int64_t j = 0; j < tSize; ++j) f.tt.ptr[j + 1] = f.tt.ptr[j] + tcount[j + 1];
  f.tt.idx.assign((size_t)f.tt.ptr[tSize], 0);
  {
    std::vector<int32_t> fill(f.tt.ptr.begin(), f.tt.ptr.end() - 1);
    for (int64_t i = 0; i < tSize; ++i)
      for (int32_t e = f.H.ptr[i]; e < f.H.ptr[i + 1]; ++e) {
        const int32_t l = f.H.idx[e];
        if (l >= N - tSize) f.tt.idx[(size_t)fill[l - (N - tSize)]++] = (int32_t)i;
      }
  }
  for (int64_t d = 0; d < P; ++d) {
    for (int64_t w = 0; w < W; ++w)
      for (uint64_t x = cde[d][w]; x; x &= x - 1) {
        const int32_t l = orig2lab[w * 64 + __builtin_ctzll(x)];
        if (l < K) tmp.push_back(l);
      }
    append_row(&f.dc, tmp);
  }
  *out = std::move(f);
  return "";
}

void encode_generator(const GeneratorForm& g, const uint8_t* msg, uint8_t* cw) {
  const int64_t K = g.K, P = g.P, t = g.t;
  for (int64_t k = 0; k < K; ++k) cw[k] = msg[k] & 1;
  std::fill(cw + K, cw + g.N, 0);
  uint8_t* p = cw + K;  // gap parity, then triangular parity
  for (int64_t d = 0; d < P; ++d) {
    uint8_t s = 0;
    for (int32_t e = g.dc.ptr[d]; e < g.dc.ptr[d + 1]; ++e) s ^= msg[g.dc.idx[e]] & 1;
    p[d] ^= s;
  }
  for (int64_t i = 0; i < t; ++i) {
    uint8_t s = 0;
    for (int32_t e = g.b.ptr[i]; e < g.b.ptr[i + 1]; ++e) s ^= p[g.b.idx[e]];
    p[P + i] ^= s;
  }
  uint8_t* q = p + P;
  for (int64_t i = 0; i < t; ++i) {
    uint8_t s = 0;
    for (int32_t e = g.a.ptr[i]; e < g.a.ptr[i + 1]; ++e) s ^= msg[g.a.idx[e]] & 1;
    q[i] ^= s;
  }
  // forward substitution through T; the first entry of each column is skipped as FeCl does
  // (it is the diagonal)
  for (int64_t j = 0; j < t; ++j)
    for (int32_t e = g.tt.ptr[j] + 1; e < g.tt.ptr[j + 1]; ++e) q[g.tt.idx[e]] ^= q[j];
}

std::string gallager_matrix(int64_t n, int64_t wc, int64_t wr, uint64_t seed, Csr* out) {
  if (n < 1 || wc < 1 || wr < 1 || wr > n) return "Gallager matrix needs 1 <= wr <= n and wc >= 1";
  if (n > (int64_t)INT32_MAX / 2 || (n / wr) * wc * wr > (int64_t)INT32_MAX / 2) return "Gallager matrix too large";
  const int64_t per = n / wr;
  // block b (rows b*per .. b*per+per-1): row i holds the columns that started as i*wr .. i*wr+wr-1,
  // relabelled by the block's column swaps; lab[b][c] is the current label of start column c
  std::vector<std::vector<int32_t>> lab((size_t)wc), inv((size_t)wc);
  for (int64_t b = 1; b < wc; ++b) {
    lab[b].resize((size_t)n);
    std::iota(lab[b].begin(), lab[b].end(), 0);
    inv[b] = lab[b];
  }
  std::minstd_rand0 generator((int)seed);
  for (int64_t j = 0; j < n; ++j) {
    std::uniform_int_distribution<int> distribution(int(j), int(n - 1));
    for (int64_t b = 1; b < wc; ++b) {
      const int64_t r = distribution(generator);
      if (r == j) continue;
      const int32_t cj = inv[b][j], cr = inv[b][r];
      std::swap(inv[b][j], inv[b][r]);
      lab[b][cj] = (int32_t)r;
      lab[b][cr] = (int32_t)j;
    }
  }
  Csr m;
  std::vector<int32_t> tmp;
  for (int64_t b = 0; b < wc; ++b)
    for (int64_t i = 0; i < per; ++i) {
      for (int64_t k = 0; k < wr; ++k) tmp.push_back(b == 0 ? (int32_t)(i * wr + k) : lab[b][i * wr + k]);
      append_row(&m, tmp);
    }
  *out = std::move(m);
  return "";
}

}  // namespace fecl
