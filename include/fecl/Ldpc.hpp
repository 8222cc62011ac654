// fecl/Ldpc.hpp — C++ host API mirroring FeCl's fec::Ldpc over the HIP decoder's C ABI.
//
// Header-only: link against fecl_amd/libfecl_ldpc.so.  Names, signatures and error behaviour
// follow the reference so existing FeCl callers compile unchanged for the LDPC decode path:
//   fec::DecoderAlgorithm                 src/DecoderAlgorithm.h:30-34
//   fec::BitField<T> (value use)          src/BitField.h:62-124
//   fec::Codec::Input / Output            src/Codec.h:51-54, src/detail/Codec.h:150-189
//   fec::Ldpc, EncoderOptions, DecoderOptions, DvbS2::matrix, Gallager::matrix
//                                         src/Ldpc.h:59-108, src/detail/Ldpc.h:44-74
//   decode / soDecode / encode / check    src/Codec.h:68-82, 134-307
//   detail::archiveSize / save / load     src/Serialization.h:81-115
// Size errors throw std::invalid_argument with FeCl's messages; device errors throw
// std::runtime_error.  workGroupSize is kept for signature parity: frames are decoded as one
// device batch whatever its value.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <type_traits>
#include <string>
#include <unordered_map>
#include <utility>
#include <algorithm>
#include <vector>

#include "../fecl_ldpc.h"

namespace fec {

enum DecoderAlgorithm { Exact = 0, Linear = 1, Approximate = 2 };

// Message storage on the device (no counterpart in FeCl, which is FP64 only).
enum class MsgFormat { F32 = FECL_MSG_F32, F16 = FECL_MSG_F16, BF16 = FECL_MSG_BF16 };

namespace detail {
inline void throw_status(int status) {
  if (status == FECL_OK) return;
  if (status == FECL_INVALID_ARGUMENT) throw std::invalid_argument(fecl_last_error());
  throw std::runtime_error(fecl_last_error());
}

template <class Vector>
struct Info {
  Info& syst(Vector& v) { syst_ = &v; return *this; }
  Info& parity(Vector& v) { parity_ = &v; return *this; }
  Info& state(Vector& v) { state_ = &v; return *this; }
  Info& msg(Vector& v) { msg_ = &v; return *this; }
  Vector& syst() const { return *syst_; }
  Vector& parity() const { return *parity_; }
  Vector& state() const { return *state_; }
  Vector& msg() const { return *msg_; }
  bool hasSyst() const { return syst_ != nullptr; }
  bool hasParity() const { return parity_ != nullptr; }
  bool hasState() const { return state_ != nullptr; }
  bool hasMsg() const { return msg_ != nullptr; }

 private:
  Vector* syst_ = nullptr;
  Vector* parity_ = nullptr;
  Vector* state_ = nullptr;
  Vector* msg_ = nullptr;
};
}  // namespace detail

// Bit holder used for hard decisions (FeCl's BitField<size_t>, value semantics only).
template <typename T>
class BitField {
 public:
  BitField() = default;
  BitField(T value) : value_(value) {}
  T operator=(T value) { value_ = value; return value_; }
  operator T() const { return value_; }
  bool test(size_t i) const { return (value_ >> i) & 1; }

 private:
  T value_ = 0;
};

// Parity-check matrix in CSR form (FeCl's SparseBitMatrix, src/BitMatrix.h:42-625): rows are
// checks, at(e) is the column of edge e in row-major order.
class SparseBitMatrix {
 public:
  SparseBitMatrix() = default;
  SparseBitMatrix(std::vector<int32_t> row_ptr, std::vector<int32_t> col, size_t cols)
      : row_ptr_(std::move(row_ptr)), col_(std::move(col)), cols_(cols) {}
  size_t rows() const { return row_ptr_.empty() ? 0 : row_ptr_.size() - 1; }
  size_t cols() const { return cols_; }
  size_t size() const { return col_.size(); }
  size_t at(size_t e) const { return (size_t)col_[e]; }
  std::vector<size_t> rowSizes() const {
    std::vector<size_t> s(rows());
    for (size_t r = 0; r < rows(); ++r) s[r] = (size_t)(row_ptr_[r + 1] - row_ptr_[r]);
    return s;
  }
  const std::vector<int32_t>& rowPtr() const { return row_ptr_; }
  const std::vector<int32_t>& colIdx() const { return col_; }
  bool operator==(const SparseBitMatrix& o) const {
    return cols_ == o.cols_ && row_ptr_ == o.row_ptr_ && col_ == o.col_;
  }

 private:
  std::vector<int32_t> row_ptr_, col_;
  size_t cols_ = 0;
};

// fec::Permutation (src/Permutation.h): output element i of a block is input element
// sequence[i]; permute maps blocks of inputSize() to outputSize(), dePermute the other way with
// the entries no index names left 0 (value-initialised, as FeCl's resized output vector).
class Permutation {
 public:
  Permutation() = default;
  explicit Permutation(const std::vector<size_t>& sequence) : sequence_(sequence) {
    if (!sequence_.empty()) inputSize_ = *std::max_element(sequence_.begin(), sequence_.end()) + 1;
  }
  Permutation(std::vector<size_t> sequence, size_t inputSize) : sequence_(std::move(sequence)), inputSize_(inputSize) {}
  size_t inputSize() const { return inputSize_; }
  size_t outputSize() const { return sequence_.size(); }
  size_t operator[](size_t i) const { return sequence_[i]; }
  template <typename T1, typename T2 = T1>
  void permute(const std::vector<T1>& input, std::vector<T2>& output) const {
    output.resize(input.size() / inputSize() * outputSize());
    for (size_t b = 0, o = 0; b < input.size(); b += inputSize(), o += outputSize())
      for (size_t i = 0; i < sequence_.size(); ++i) output[o + i] = input[b + sequence_[i]];
  }
  template <typename T1, typename T2 = T1>
  void dePermute(const std::vector<T1>& input, std::vector<T2>& output) const {
    output.assign(input.size() / outputSize() * inputSize(), T2());
    for (size_t b = 0, o = 0; b < input.size(); b += outputSize(), o += inputSize())
      for (size_t i = 0; i < sequence_.size(); ++i) output[o + sequence_[i]] = input[b + i];
  }
  template <typename T>
  std::vector<T> permute(const std::vector<T>& input) const { std::vector<T> o; permute(input, o); return o; }
  template <typename T>
  std::vector<T> dePermute(const std::vector<T>& input) const { std::vector<T> o; dePermute(input, o); return o; }

 private:
  std::vector<size_t> sequence_;
  size_t inputSize_ = 0;
};

class Codec {
 public:
  template <template <typename> class A = std::allocator>
  using Input = detail::Info<const std::vector<double, A<double>>>;
  template <template <typename> class A = std::allocator>
  using Output = detail::Info<std::vector<double, A<double>>>;
};

class Ldpc : public Codec {
 public:
  // Ldpc::Gallager (src/Ldpc.h:63-65): the same matrix as FeCl for the same arguments
  struct Gallager {
    static SparseBitMatrix matrix(size_t n, size_t wc, size_t wr, uint64_t seed = 0) {
      int64_t dims[3];
      detail::throw_status(fecl_gallager_matrix((int64_t)n, (int64_t)wc, (int64_t)wr, seed, dims, nullptr, nullptr));
      std::vector<int32_t> rp(dims[0] + 1), col(dims[2]);
      detail::throw_status(fecl_gallager_matrix((int64_t)n, (int64_t)wc, (int64_t)wr, seed, dims, rp.data(), col.data()));
      return SparseBitMatrix(std::move(rp), std::move(col), (size_t)dims[1]);
    }
  };
  struct DvbS2 {
    static SparseBitMatrix matrix(size_t n, double rate) {
      static const int rates[][2] = {{1, 4}, {1, 3}, {2, 5}, {1, 2}, {3, 5}, {2, 3},
                                     {3, 4}, {4, 5}, {5, 6}, {8, 9}, {9, 10}};
      if (n != 64800 && n != 16200) throw std::invalid_argument("Invalid length");
      for (const auto& r : rates) {
        if (rate != (double)r[0] / (double)r[1]) continue;
        int64_t dims[4];
        if (fecl_dvbs2_matrix((int32_t)n, r[0], r[1], dims, nullptr, nullptr) != FECL_OK) break;
        std::vector<int32_t> rp(dims[0] + 1), col(dims[3]);
        detail::throw_status(fecl_dvbs2_matrix((int32_t)n, r[0], r[1], dims, rp.data(), col.data()));
        return SparseBitMatrix(std::move(rp), std::move(col), (size_t)dims[1]);
      }
      throw std::invalid_argument("Invalid rate");
    }
  };

  struct EncoderOptions {
    EncoderOptions(const SparseBitMatrix& checkMatrix) : checkMatrix_(checkMatrix) {}
    SparseBitMatrix checkMatrix() const { return checkMatrix_; }
    SparseBitMatrix checkMatrix_;
  };

  // Ldpc::PunctureOptions (src/detail/Ldpc.h:76-82)
  struct PunctureOptions {
    PunctureOptions() = default;
    PunctureOptions& mask(std::vector<bool> m) { mask_ = std::move(m); return *this; }
    PunctureOptions& systMask(std::vector<bool> m) { systMask_ = std::move(m); return *this; }
    std::vector<bool> mask_, systMask_;
  };

  struct DecoderOptions {
    DecoderOptions() = default;
    DecoderOptions& algorithm(DecoderAlgorithm a) { algorithm_ = a; return *this; }
    DecoderOptions& iterations(size_t n) { iterations_ = n; return *this; }
    DecoderOptions& scalingFactor(double f) { scalingFactor_ = {{0, std::vector<double>{f}}}; return *this; }
    DecoderOptions& scalingFactor(const std::unordered_map<size_t, std::vector<double>>& f) {
      scalingFactor_ = f;
      return *this;
    }
    DecoderOptions& msgFormat(MsgFormat f) { msgFormat_ = f; return *this; }
    DecoderAlgorithm algorithm() const { return algorithm_; }
    size_t iterations() const { return iterations_; }
    std::unordered_map<size_t, std::vector<double>> scalingFactor() const { return scalingFactor_; }
    MsgFormat msgFormat() const { return msgFormat_; }

    DecoderAlgorithm algorithm_ = Approximate;
    size_t iterations_ = 0;  // FeCl leaves it uninitialised; 0 is rejected on use
    std::unordered_map<size_t, std::vector<double>> scalingFactor_ = {{0, std::vector<double>{1.0}}};
    MsgFormat msgFormat_ = MsgFormat::F32;
  };

  Ldpc(const EncoderOptions& encoder, const DecoderOptions& decoder, int workGroupSize = 8)
      : H_(encoder.checkMatrix_), workGroupSize_(workGroupSize) {
    std::vector<double> table;
    fecl_decoder_options o = options(decoder, &table);
    detail::throw_status(fecl_ldpc_create(H_.rowPtr().data(), H_.colIdx().data(), (int64_t)H_.rows(),
                                          (int64_t)H_.cols(), &o, &h_));
    decoder_ = decoder;
    detail::throw_status(fecl_ldpc_sizes(h_, &msg_, &syst_, &parity_, &state_));
  }
  Ldpc(const Ldpc&) = delete;
  Ldpc& operator=(const Ldpc&) = delete;
  Ldpc(Ldpc&& o) noexcept { *this = std::move(o); }
  Ldpc& operator=(Ldpc&& o) noexcept {
    std::swap(h_, o.h_);
    H_ = o.H_;
    decoder_ = o.decoder_;
    workGroupSize_ = o.workGroupSize_;
    msg_ = o.msg_; syst_ = o.syst_; parity_ = o.parity_; state_ = o.state_;
    return *this;
  }
  ~Ldpc() {
    if (h_) fecl_ldpc_destroy(h_);
  }

  size_t msgSize() const { return (size_t)msg_; }
  size_t systSize() const { return (size_t)syst_; }
  size_t paritySize() const { return (size_t)parity_; }
  size_t stateSize() const { return (size_t)state_; }
  int getWorkGroupSize() const { return workGroupSize_; }
  void setWorkGroupSize(int size) { workGroupSize_ = size; }

  void setDecoderOptions(const DecoderOptions& decoder) {
    std::vector<double> table;
    fecl_decoder_options o = options(decoder, &table);
    detail::throw_status(fecl_ldpc_set_options(h_, &o));
    decoder_ = decoder;
  }
  DecoderOptions getDecoderOptions() const { return decoder_; }

  // fec::detail::save / load (src/Serialization.h:81-115) of this codec: FeCl's archive bytes
  // (Boost binary archive, the Matlab wrapper's type registration).  load() keeps the stored
  // structure as it is and uses FP32 messages.
  std::vector<char> save() const {
    int64_t n = 0;
    detail::throw_status(fecl_ldpc_save(h_, workGroupSize_, nullptr, 0, &n));
    std::vector<char> out((size_t)n);
    detail::throw_status(fecl_ldpc_save(h_, workGroupSize_, reinterpret_cast<uint8_t*>(out.data()), n, &n));
    return out;
  }
  static Ldpc load(const char* buffer, size_t size) {
    fecl_ldpc* h = nullptr;
    int32_t wgs = 8;
    detail::throw_status(fecl_ldpc_load(reinterpret_cast<const uint8_t*>(buffer), (int64_t)size, &h, &wgs));
    return Ldpc(h, wgs);
  }

  // Ldpc::puncturing (src/Ldpc.h:94; index rule of src/detail/Ldpc.cpp:333-353)
  Permutation puncturing(const PunctureOptions& options) const {
    std::vector<uint8_t> m(options.mask_.begin(), options.mask_.end());
    std::vector<uint8_t> sm(options.systMask_.begin(), options.systMask_.end());
    int64_t n = 0;
    detail::throw_status(fecl_ldpc_puncturing(syst_, parity_, m.data(), (int64_t)m.size(), sm.data(),
                                              (int64_t)sm.size(), nullptr, 0, &n));
    std::vector<int64_t> seq((size_t)std::max<int64_t>(n, 1));
    detail::throw_status(fecl_ldpc_puncturing(syst_, parity_, m.data(), (int64_t)m.size(), sm.data(),
                                              (int64_t)sm.size(), seq.data(), (int64_t)seq.size(), &n));
    return Permutation(std::vector<size_t>(seq.begin(), seq.begin() + n), (size_t)parity_);
  }

  template <template <typename> class A, class B>
  bool check(const std::vector<B, A<B>>& parity) const {
    const size_t frames = parity.size() / paritySize();
    if (parity.size() != frames * paritySize()) throw std::invalid_argument("Invalid size for parity");
    std::vector<uint8_t> p(parity.begin(), parity.end());
    int32_t ok = 0;
    detail::throw_status(fecl_ldpc_check(h_, p.data(), (int64_t)frames, &ok));
    return ok != 0;
  }
  template <class B>
  bool check(const std::vector<B>& parity) const { return check<std::allocator, B>(parity); }

  template <template <typename> class A, class B>
  void encode(const std::vector<B, A<B>>& msg, std::vector<B, A<B>>& parity) const {
    const size_t frames = msg.size() / msgSize();
    if (msg.size() != frames * msgSize()) throw std::invalid_argument("Invalid size for message");
    std::vector<uint8_t> m(msg.size()), p(frames * paritySize());
    for (size_t i = 0; i < msg.size(); ++i) m[i] = (uint8_t)(msg[i] ? 1 : 0);
    detail::throw_status(fecl_ldpc_encode(h_, m.data(), (int64_t)frames, p.data()));
    parity.assign(p.begin(), p.end());
  }
  template <class B>
  void encode(const std::vector<B>& msg, std::vector<B>& parity) const { encode<std::allocator, B>(msg, parity); }
  template <class B>
  std::vector<B> encode(const std::vector<B>& msg) const {
    std::vector<B> parity;
    encode(msg, parity);
    return parity;
  }

  // Codec::decode: parity L-values [frames][N] -> msg bits [frames][K].
  template <template <typename> class A, class B>
  void decode(const std::vector<double, A<double>>& parity, std::vector<B, A<B>>& msg) const {
    const size_t frames = parity.size() / paritySize();
    if (parity.size() != frames * paritySize()) throw std::invalid_argument("Invalid size for parity");
    std::vector<uint8_t> bits(frames * msgSize());
    detail::throw_status(fecl_ldpc_decode(h_, parity.data(), (int64_t)frames, bits.data(), nullptr));
    msg.assign(bits.begin(), bits.end());
  }
  template <class B>
  void decode(const std::vector<double>& parity, std::vector<B>& msg) const { decode<std::allocator, B>(parity, msg); }
  std::vector<BitField<size_t>> decode(const std::vector<double>& parity) const {
    std::vector<BitField<size_t>> msg;
    decode(parity, msg);
    return msg;
  }

  // Codec::soDecode with FeCl's validation order and messages (src/Codec.h:252-270).
  template <template <typename> class A = std::allocator>
  void soDecode(Input<A> input, Output<A> output) const {
    if (!input.hasParity()) throw std::invalid_argument("Input must contains parity");
    if (input.hasMsg()) throw std::invalid_argument("Input should not contain msg");
    const size_t frames = input.parity().size() / paritySize();
    if (input.parity().size() != frames * paritySize()) throw std::invalid_argument("Invalid size for parity");
    if (input.hasSyst() && input.syst().size() != frames * systSize()) throw std::invalid_argument("Invalid size for msg");
    if (input.hasState() && input.state().size() != frames * stateSize())
      throw std::invalid_argument("Invalid size for state");
    if (output.hasParity()) output.parity().resize(frames * paritySize());
    if (output.hasSyst()) output.syst().resize(frames * systSize());
    if (output.hasState()) output.state().resize(frames * stateSize());
    if (output.hasMsg()) output.msg().resize(frames * msgSize());
    detail::throw_status(fecl_ldpc_sodecode(
        h_, input.parity().data(), input.hasSyst() ? input.syst().data() : nullptr,
        input.hasState() ? input.state().data() : nullptr, (int64_t)frames,
        output.hasMsg() ? output.msg().data() : nullptr, output.hasSyst() ? output.syst().data() : nullptr,
        output.hasParity() ? output.parity().data() : nullptr, output.hasState() ? output.state().data() : nullptr,
        nullptr));
  }

  // Per-frame check-node update counts of a decode (FeCl computes but does not expose them).
  std::vector<int32_t> decodeIterations(const std::vector<double>& parity, std::vector<uint8_t>* bits = nullptr) const {
    const size_t frames = parity.size() / paritySize();
    if (parity.size() != frames * paritySize()) throw std::invalid_argument("Invalid size for parity");
    std::vector<uint8_t> b(frames * msgSize());
    std::vector<int32_t> it(frames);
    detail::throw_status(fecl_ldpc_decode(h_, parity.data(), (int64_t)frames, b.data(), it.data()));
    if (bits) *bits = std::move(b);
    return it;
  }

  fecl_ldpc* handle() const { return h_; }

 private:
  // a codec around a handle made by fecl_ldpc_load: its matrix and options come from the handle
  Ldpc(fecl_ldpc* h, int workGroupSize) : workGroupSize_(workGroupSize), h_(h) {
    int64_t dims[4];
    detail::throw_status(fecl_ldpc_checks(h_, dims, nullptr, nullptr));
    std::vector<int32_t> rp(dims[0] + 1), col(dims[3]);
    detail::throw_status(fecl_ldpc_checks(h_, dims, rp.data(), col.data()));
    H_ = SparseBitMatrix(std::move(rp), std::move(col), (size_t)dims[1]);
    fecl_decoder_options o{};
    detail::throw_status(fecl_ldpc_get_options(h_, &o, nullptr, 0));
    std::vector<double> table((size_t)o.scaling_rows * o.scaling_cols);
    detail::throw_status(fecl_ldpc_get_options(h_, &o, table.data(), (int64_t)table.size()));
    std::unordered_map<size_t, std::vector<double>> map;
    for (int32_t d = 0; d < o.scaling_cols; ++d)
      for (int32_t r = 0; r < o.scaling_rows; ++r) map[(size_t)d + 2].push_back(table[(size_t)r * o.scaling_cols + d]);
    decoder_ = DecoderOptions().algorithm((DecoderAlgorithm)o.algorithm).iterations((size_t)o.iterations).scalingFactor(map);
    detail::throw_status(fecl_ldpc_sizes(h_, &msg_, &syst_, &parity_, &state_));
  }

  fecl_decoder_options options(const DecoderOptions& d, std::vector<double>* table) const {
    if (d.iterations_ < 1) throw std::invalid_argument("iterations must be >= 1");
    std::vector<int64_t> keys, lengths, offsets;
    std::vector<double> values;
    for (const auto& kv : d.scalingFactor_) {
      keys.push_back((int64_t)kv.first);
      lengths.push_back((int64_t)kv.second.size());
      offsets.push_back((int64_t)values.size());
      values.insert(values.end(), kv.second.begin(), kv.second.end());
    }
    size_t maxdeg = 0;
    for (size_t s : H_.rowSizes()) maxdeg = s > maxdeg ? s : maxdeg;
    table->assign(d.iterations_ * (maxdeg > 1 ? maxdeg - 1 : 1), 0.0);
    int32_t rows = 0, cols = 0;
    detail::throw_status(fecl_ldpc_scaling_table(H_.rowPtr().data(), (int64_t)H_.rows(), (int64_t)d.iterations_,
                                                 keys.data(), lengths.data(), offsets.data(), values.data(),
                                                 (int64_t)keys.size(), table->data(), (int64_t)table->size(),
                                                 &rows, &cols));
    fecl_decoder_options o{};
    o.algorithm = (int32_t)d.algorithm_;
    o.iterations = (int32_t)d.iterations_;
    o.scaling = table->data();
    o.scaling_rows = rows;
    o.scaling_cols = cols;
    o.msg_format = (int32_t)d.msgFormat_;
    return o;
  }

  SparseBitMatrix H_;
  DecoderOptions decoder_;
  int workGroupSize_ = 8;
  fecl_ldpc* h_ = nullptr;
  int64_t msg_ = 0, syst_ = 0, parity_ = 0, state_ = 0;
};

namespace detail {
// fec::detail::archiveSize / save / load (src/Serialization.h:81-115) for Ldpc codecs.  The
// size is exact (FeCl's adds 4096 bytes of slack); load() also accepts buffers with slack.
inline size_t archiveSize(const Ldpc* codec) { return codec->save().size(); }
inline void save(const Ldpc* codec, char* buffer, size_t size) {
  std::vector<char> bytes = codec->save();
  if (bytes.size() > size) throw std::invalid_argument("archive buffer too small");
  std::copy(bytes.begin(), bytes.end(), buffer);
}
template <class T = Ldpc>
std::unique_ptr<T> load(const char* buffer, size_t size) {
  static_assert(std::is_same<T, Ldpc>::value, "saved codecs load as fec::Ldpc");
  return std::unique_ptr<T>(new Ldpc(Ldpc::load(buffer, size)));
}
}  // namespace detail

}  // namespace fec
