#include "ldpc_archive.h"

#include <climits>
#include <cstring>
#include <stdexcept>

namespace fecl {
namespace {

// Boost.Serialization binary archive of library version 12 (Boost 1.58), little-endian x86-64:
// strings and collection sizes as 8-byte counts, class ids as int16, object ids and versions
// as uint32, tracking flags as one byte.  A class's tracking flag and version precede its first
// object in the archive only; pointers add the class id and an object id.
constexpr char kSignature[] = "serialization::archive";
constexpr uint16_t kLibraryVersion = 12;
constexpr int16_t kLdpcClass = 2;        // third type registered (Convolutional, Turbo, Ldpc)
constexpr int16_t kStructureClass = 3;   // registered by Ldpc::serialize (src/Ldpc.h:121)

struct Writer {
  std::vector<uint8_t>* b;
  template <typename T> void put(T v) {
    uint8_t x[sizeof(T)];
    std::memcpy(x, &v, sizeof(T));
    b->insert(b->end(), x, x + sizeof(T));
  }
  void class_info() {   // tracking off, version 0
    put<uint8_t>(0);
    put<uint32_t>(0);
  }
};

struct Reader {
  const uint8_t* p;
  int64_t left;
  template <typename T> T get() {
    if (left < (int64_t)sizeof(T)) throw std::runtime_error("archive is truncated");
    T v;
    std::memcpy(&v, p, sizeof(T));
    p += sizeof(T);
    left -= sizeof(T);
    return v;
  }
  template <typename T> void expect(T want, const char* what) {
    if (get<T>() != want) throw std::runtime_error(std::string("unexpected ") + what + " in archive");
  }
  void class_info(const char* what) {
    get<uint8_t>();   // tracking flag: informational here
    expect<uint32_t>(0, what);
  }
  uint64_t count(int64_t elem_bytes) {
    uint64_t n = get<uint64_t>();
    if (elem_bytes > 0 && n > (uint64_t)(left / elem_bytes)) throw std::runtime_error("archive is truncated");
    return n;
  }
};

// SparseBitMatrix::serialize (src/BitMatrix.h:614-620): cols_, elementIdx_ (vector<size_t>,
// written as a raw array), rowIdx_ (vector<RowIdx{begin, end}>, with an item version).  The
// matrices here are compact: row r spans [ptr[r], ptr[r+1]).
void put_matrix(Writer& w, const Csr& m, int64_t cols, bool first) {
  if (first) w.class_info();                       // SparseBitMatrix
  w.put<uint64_t>((uint64_t)cols);
  w.put<uint64_t>((uint64_t)m.idx.size());
  for (int32_t c : m.idx) w.put<uint64_t>((uint64_t)c);
  if (first) w.class_info();                       // std::vector<RowIdx>
  w.put<uint64_t>((uint64_t)m.rows());
  w.put<uint32_t>(0);                              // item version
  if (first && m.rows() > 0) w.class_info();       // RowIdx, before its first object
  for (int64_t r = 0; r < m.rows(); ++r) {
    w.put<uint64_t>((uint64_t)m.ptr[r]);
    w.put<uint64_t>((uint64_t)m.ptr[r + 1]);
  }
}

Csr get_matrix(Reader& r, int64_t* cols, bool first, bool* rowidx_seen) {
  if (first) r.class_info("SparseBitMatrix version");
  const uint64_t c = r.get<uint64_t>();
  if (c > (uint64_t)INT32_MAX) throw std::runtime_error("matrix too large for this library");
  *cols = (int64_t)c;
  const uint64_t nnz = r.count(8);
  if (nnz > (uint64_t)INT32_MAX) throw std::runtime_error("matrix too large for this library");
  std::vector<uint64_t> el(nnz);
  for (auto& e : el) e = r.get<uint64_t>();
  if (first) r.class_info("row index version");
  const uint64_t rows = r.count(16);
  r.expect<uint32_t>(0, "item version");
  if (!*rowidx_seen && rows > 0) {
    r.class_info("row index version");
    *rowidx_seen = true;
  }
  Csr m;
  // FeCl's rows are consecutive ranges of the element list ({begin, end} with begin = the
  // previous row's end, the first at 0), so the expanded matrix never exceeds nnz entries
  uint64_t next = 0;
  for (uint64_t i = 0; i < rows; ++i) {
    const uint64_t b = r.get<uint64_t>(), e = r.get<uint64_t>();
    if (b > e || e > nnz) throw std::runtime_error("row index outside the matrix");
    if (b != next) throw std::runtime_error("row ranges are not consecutive");
    next = e;
    for (uint64_t k = b; k < e; ++k) {
      if (el[k] >= (uint64_t)*cols) throw std::runtime_error("column index outside the matrix");
      m.idx.push_back((int32_t)el[k]);
    }
    m.ptr.push_back((int32_t)m.idx.size());
  }
  return m;
}

}  // namespace

void archive_write(const GeneratorForm& g, int64_t E, const ArchiveOptions& o, std::vector<uint8_t>* out) {
  Writer w{out};
  w.put<uint64_t>(sizeof(kSignature) - 1);
  out->insert(out->end(), kSignature, kSignature + sizeof(kSignature) - 1);
  w.put<uint16_t>(kLibraryVersion);
  for (uint8_t s : {4, 8, 4, 8}) w.put<uint8_t>(s);   // sizeof int, long, float, double
  w.put<int32_t>(1);                                 // byte-order check
  // const fec::Codec* -> fec::Ldpc: class id, tracked, version 0, object 0
  w.put<int16_t>(kLdpcClass);
  w.put<uint8_t>(1);
  w.put<uint32_t>(0);
  w.put<uint32_t>(0);
  w.class_info();                                    // base fec::Codec
  w.put<int32_t>(o.work_group_size);
  w.class_info();                                    // std::unique_ptr<Codec::Structure>
  // the structure pointer: Ldpc::Structure, tracked, version 0, object 1
  w.put<int16_t>(kStructureClass);
  w.put<uint8_t>(1);
  w.put<uint32_t>(0);
  w.put<uint32_t>(1);
  w.class_info();                                    // base Codec::Structure
  w.put<uint64_t>((uint64_t)g.K);                    // msgSize_
  w.put<uint64_t>((uint64_t)g.K);                    // systSize_
  w.put<uint64_t>((uint64_t)g.N);                    // paritySize_
  w.put<uint64_t>((uint64_t)E);                      // stateSize_
  w.put<int32_t>(o.algorithm);                       // decoderAlgorithm_
  put_matrix(w, g.H, g.N, true);
  put_matrix(w, g.dc, g.K, false);
  put_matrix(w, g.tt, g.t, false);
  put_matrix(w, g.a, g.K, false);
  put_matrix(w, g.b, g.P, false);
  w.put<uint64_t>((uint64_t)o.iterations);
  w.class_info();                                    // std::vector<std::vector<double>>
  w.put<uint64_t>((uint64_t)o.rows);
  w.put<uint32_t>(0);                                // item version
  for (int64_t r = 0; r < o.rows; ++r) {
    w.put<uint64_t>((uint64_t)o.cols);
    for (int64_t c = 0; c < o.cols; ++c) w.put<double>(o.scaling[(size_t)(r * o.cols + c)]);
  }
}

std::string archive_read(const uint8_t* buf, int64_t size, GeneratorForm* g, ArchiveOptions* o) {
  if (!buf || size <= 0) return "empty archive";
  try {
    Reader r{buf, size};
    const uint64_t sl = r.get<uint64_t>();
    if (sl != sizeof(kSignature) - 1 || r.left < (int64_t)sl || std::memcmp(r.p, kSignature, sl) != 0)
      return "not a Boost binary archive";
    r.p += sl;
    r.left -= (int64_t)sl;
    if (r.get<uint16_t>() != kLibraryVersion) return "archive library version is not 12 (Boost 1.58)";
    for (uint8_t s : {4, 8, 4, 8}) r.expect<uint8_t>(s, "primitive size");
    r.expect<int32_t>(1, "byte order");
    if (r.get<int16_t>() != kLdpcClass) return "archive does not hold an Ldpc codec";
    r.get<uint8_t>();
    r.expect<uint32_t>(0, "Ldpc version");
    r.expect<uint32_t>(0, "object id");
    r.class_info("Codec version");
    ArchiveOptions opt;
    opt.work_group_size = r.get<int32_t>();
    r.class_info("pointer holder version");
    if (r.get<int16_t>() != kStructureClass) return "archive does not hold an Ldpc structure";
    r.get<uint8_t>();
    r.expect<uint32_t>(0, "Structure version");
    r.expect<uint32_t>(1, "object id");
    r.class_info("Codec::Structure version");
    GeneratorForm f;
    const uint64_t msg = r.get<uint64_t>(), syst = r.get<uint64_t>(), parity = r.get<uint64_t>(),
                   state = r.get<uint64_t>();
    opt.algorithm = r.get<int32_t>();
    bool rowidx_seen = false;
    int64_t hc = 0, dcc = 0, tc = 0, ac = 0, bc = 0;
    f.H = get_matrix(r, &hc, true, &rowidx_seen);
    f.dc = get_matrix(r, &dcc, false, &rowidx_seen);
    f.tt = get_matrix(r, &tc, false, &rowidx_seen);
    f.a = get_matrix(r, &ac, false, &rowidx_seen);
    f.b = get_matrix(r, &bc, false, &rowidx_seen);
    opt.iterations = (int64_t)r.get<uint64_t>();
    r.class_info("scaling table version");
    opt.rows = (int64_t)r.count(8);
    r.expect<uint32_t>(0, "item version");
    for (int64_t i = 0; i < opt.rows; ++i) {
      const int64_t n = (int64_t)r.count(8);
      if (i == 0) opt.cols = n;
      if (n != opt.cols) return "scaling table rows differ in length";
      for (int64_t c = 0; c < n; ++c) opt.scaling.push_back(r.get<double>());
    }
    f.M = f.H.rows();
    f.N = hc;
    f.K = (int64_t)msg;
    f.t = f.tt.rows();
    f.P = f.dc.rows();
    if (syst != msg || parity != (uint64_t)f.N || state != f.H.idx.size())
      return "archive sizes do not match its parity-check matrix";
    if (f.M < 1 || f.N <= f.M || f.K < f.N - f.M || f.K >= f.N || f.K + f.P + f.t != f.N || dcc != f.K ||
        tc != f.t || ac != f.K || bc != f.P || f.a.rows() != f.t || f.b.rows() != f.t)
      return "archive encoder blocks do not match its parity-check matrix";
    if (opt.algorithm < 0 || opt.algorithm > 2) return "archive holds an unknown decoder algorithm";
    *g = std::move(f);
    *o = std::move(opt);
    return "";
  } catch (const std::exception& e) {
    return e.what();
  }
}

}  // namespace fecl
