#include "ldpc_code.h"

#include <algorithm>

#include "dvbs2_tables.h"

namespace fecl {

std::string build_code(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, Code* out) {
  if (!row_ptr || !col) return "null parity-check arrays";
  if (M < 1 || N < 1 || M >= N) return "parity-check matrix must have 1 <= rows < cols";
  if (row_ptr[0] != 0) return "row_ptr[0] must be 0";
  Code c;
  c.M = M;
  c.N = N;
  c.K = N - M;
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

void encode_staircase(const Code& c, const uint8_t* msg, uint8_t* cw) {
  std::copy(msg, msg + c.K, cw);
  uint8_t acc = 0;
  for (int64_t r = 0; r < c.M; ++r) {
    uint8_t s = 0;
    for (int32_t e = c.row_ptr[r]; e < c.row_ptr[r + 1]; ++e)
      if (c.col[e] < c.K) s ^= msg[c.col[e]] & 1;
    acc ^= s;
    cw[c.K + r] = acc;
  }
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

}  // namespace fecl
