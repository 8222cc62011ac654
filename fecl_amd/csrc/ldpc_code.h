// Host-side parity-check structures for the LDPC decoder.
//
// A code is its parity-check matrix H in CSR form: rows are checks, and the edges of row r are
// row_ptr[r] .. row_ptr[r+1]-1 in that order.  This is FeCl's SparseBitMatrix layout
// (src/BitMatrix.h:623-624, elementIdx_ = column of edge e in row-major order), and the edge
// order fixes both the state layout (soDecode state in/out) and the variable-node summation
// order.  The column view (per-variable ascending edge lists) is derived once here.
#pragma once
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace fecl {

struct Code {
  int64_t M = 0, N = 0, K = 0, E = 0;
  std::vector<int32_t> row_ptr;     // [M+1]
  std::vector<int32_t> col;         // [E]   variable of edge e
  std::vector<int32_t> var_ptr;     // [N+1]
  std::vector<int32_t> var_edges;   // [E]   edges of variable v, ascending
  int32_t max_row_degree = 0;
  int32_t max_var_degree = 0;
  bool staircase = false;           // H = [A | dual-diagonal], so the systematic encoder applies
};

// Validates and completes a code from CSR arrays.  K = N - M (src/detail/Ldpc.cpp:46-55).
// Returns an empty string on success, otherwise the reason.
std::string build_code(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, Code* out);

// DVB-S2 parity-check matrix (reference: Ldpc::DvbS2::matrix, src/DvbS2.cpp:1595-1645), built
// for every normal (n = 64800) and short (n = 16200) rate of ETSI EN 302 307-1.
// dims = {M, N, K, E}; row_ptr/col may be null to query the sizes.  Returns false if unknown.
bool dvbs2_matrix(int n, int rate_num, int rate_den, int64_t* dims, int32_t* row_ptr, int32_t* col);

// Systematic encoder for staircase codes: codeword = [msg | parity] with
// p_r = p_{r-1} xor (xor of the message bits on check r).
void encode_staircase(const Code& c, const uint8_t* msg, uint8_t* cw);

// FeCl's scaling-factor rules (Structure::scalingMapToVector, src/detail/Ldpc.cpp:64-112):
// a map {row degree -> per-iteration factors} with key 0 as the default becomes a table
// [rows][max_degree-1] with rows = 1 or iterations.  Returns "" or FeCl's error message.
std::string scaling_map_to_table(const std::unordered_map<size_t, std::vector<double>>& map,
                                 const Code& c, int64_t iterations,
                                 std::vector<double>* table, int32_t* rows, int32_t* cols);

}  // namespace fecl
