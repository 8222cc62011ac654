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

// Validates and completes a code from CSR arrays.  K = N - M unless given (k >= 0): a code
// whose checks are linearly dependent carries more message bits (src/detail/Ldpc.cpp:307).
// Returns an empty string on success, otherwise the reason.
std::string build_code(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N, Code* out,
                       int64_t k = -1);

// A sparse 0/1 matrix as CSR: row r holds idx[ptr[r] .. ptr[r+1]), ascending.
struct Csr {
  std::vector<int32_t> ptr{0};
  std::vector<int32_t> idx;
  int64_t rows() const { return (int64_t)ptr.size() - 1; }
};

// The form FeCl puts a parity-check matrix in before decoding or encoding
// (Structure::computeGeneratorMatrix, src/detail/Ldpc.cpp:216-331).  Its columns and rows are
// permuted so that, with t triangular rows and P = N - K - t gap columns,
//   H_ = [ A | B | T ]   rows [0, t): T (the last t columns) is lower triangular
//        [ C | D | E ]   rows [t, M): [C | D] reduces to [DC | I] once T is eliminated
// and a codeword is [message (K) | gap parity (P) | triangular parity (t)].  The decoder runs on
// H_, so message bits, state order and iteration order all follow it.
struct GeneratorForm {
  int64_t M = 0, N = 0, K = 0, t = 0, P = 0;
  Csr H;       // H_ itself (rows M, columns N)
  Csr dc;      // P rows over message columns: gap parity p_d = xor of msg[dc row d]
  Csr a;       // t rows over message columns
  Csr b;       // t rows over gap parity columns (0 .. P-1)
  Csr tt;      // T transposed: row j lists the triangular rows holding column N-t+j
};

// Puts H (CSR, columns ascending within each row) in FeCl's generator form.  Returns "" or the
// reason the matrix is rejected.
std::string generator_form(const int32_t* row_ptr, const int32_t* col, int64_t M, int64_t N,
                           GeneratorForm* out);

// Structure::encode (src/detail/Ldpc.cpp:180-209) on a generator form: cw[0..N) from msg[0..K).
void encode_generator(const GeneratorForm& g, const uint8_t* msg, uint8_t* cw);

// Ldpc::Gallager::matrix (src/Ldpc.cpp:77-100): wc stacked blocks of n/wr rows of wr ones;
// block 0 in order, the columns of every other block shuffled by a minstd_rand0 generator
// seeded with (int)seed.  Returns "" or the reason the arguments are rejected.
std::string gallager_matrix(int64_t n, int64_t wc, int64_t wr, uint64_t seed, Csr* out);

// DVB-S2 parity-check matrix (reference: Ldpc::DvbS2::matrix, src/DvbS2.cpp:1595-1645), built
// for every normal (n = 64800) and short (n = 16200) rate of ETSI EN 302 307-1.
// dims = {M, N, K, E}; row_ptr/col may be null to query the sizes.  Returns false if unknown.
bool dvbs2_matrix(int n, int rate_num, int rate_den, int64_t* dims, int32_t* row_ptr, int32_t* col);

// FeCl's scaling-factor rules (Structure::scalingMapToVector, src/detail/Ldpc.cpp:64-112):
// a map {row degree -> per-iteration factors} with key 0 as the default becomes a table
// [rows][max_degree-1] with rows = 1 or iterations.  Returns "" or FeCl's error message.
std::string scaling_map_to_table(const std::unordered_map<size_t, std::vector<double>>& map,
                                 const Code& c, int64_t iterations,
                                 std::vector<double>* table, int32_t* rows, int32_t* cols);

}  // namespace fecl
