// MATLAB gateway for the LDPC codec: FeCl's `fec.bin.wrap` MEX function (wrap/Matlab/wrap.cpp,
// wrap/wrap.h:48-247) re-implemented over this library's C++ API (include/fecl/Ldpc.hpp), so
// FeCl's own MATLAB classes (+fec/Codec.m, +fec/Ldpc.m, ...) run the decoder on the GPU.
//
// The function ids are FeCl's (+fec/+detail/WrapFcnId.m), and so are the argument and result
// conventions (wrap/Matlab/util/Conversion.h, BitMatrix.h, Permutation.h, SaveLoad.h,
// Ldpc/*.h):
//   * a codec object carries its handle, a uint64 pointer, in the property `mexHandle_`;
//   * numeric inputs of any real class are converted; bit vectors come back as uint64 columns,
//     LLRs as double columns, sizes as uint64 scalars, the work group size as int32;
//   * parity-check matrices are sparse (or dense) M x N arrays; DvbS2 matrices come back as
//     sparse logical;
//   * decoder options are a struct {iterations, algorithm, scalingFactor}, with the scaling map
//     a struct array of {key, value};  puncture options a struct {mask, systMask};
//   * permutations are 1-based index columns; saved codecs are uint8 columns in FeCl's archive
//     format, padded as FeCl's save pads them.
// Turbo, Convolutional and Trellis ids report an error: only LDPC codecs are provided.
//
// Build (MATLAB):  mex -R2017b -I<repo>/include fecl_wrap.cpp -L<repo>/fecl_amd -lfecl_ldpc -output wrap
// and put the result where FeCl's classes call it (+fec/+bin/).
#include <mex.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "fecl/Ldpc.hpp"

namespace {

// FeCl's function ids (+fec/+detail/WrapFcnId.m)
enum FcnId : size_t {
  Codec_destroy = 0, Codec_save, Codec_load, Codec_check, Codec_encode, Codec_decode, Codec_soDecode,
  Codec_get_msgSize, Codec_get_systSize, Codec_get_stateSize, Codec_get_paritySize, Codec_get_workGroupSize,
  Codec_set_workGroupSize, Turbo_constructor, Turbo_get_decoderOptions, Turbo_set_decoderOptions,
  Turbo_puncturing, Turbo_Lte3Gpp_interleaver, Ldpc_constructor, Ldpc_get_decoderOptions,
  Ldpc_set_decoderOptions, Ldpc_puncturing, Ldpc_DvbS2_matrix, Convolutional_constructor,
  Convolutional_get_decoderOptions, Convolutional_set_decoderOptions, Convolutional_puncturing,
  Trellis_constructor, kFcnCount
};

// The object a handle points to; the tag rejects pointers that are not ours.
struct CodecBox {
  static constexpr uint64_t kTag = 0x4665436c4c647063ull;   // "FeClLdpc"
  uint64_t tag = kTag;
  fec::Ldpc codec;
  explicit CodecBox(fec::Ldpc&& c) : codec(std::move(c)) {}
};

struct Context {   // the "In <what>: " prefix FeCl's converters put in front of messages
  std::string what;
  [[noreturn]] void fail(const std::string& msg) const { throw std::invalid_argument(what + msg); }
};

// ---- MATLAB -> C++
template <class T>
T scalar(const mxArray* in, const Context& cx) {
  if (!in) cx.fail("Null input");
  if (mxIsComplex(in)) cx.fail("Input is complex");
  if (mxGetNumberOfElements(in) != 1 || !mxGetData(in)) cx.fail("Input is not a scalar");
  const void* d = mxGetData(in);
  switch (mxGetClassID(in)) {
    case mxLOGICAL_CLASS: return (T) * (const mxLogical*)d;
    case mxINT8_CLASS: return (T) * (const int8_t*)d;
    case mxUINT8_CLASS: return (T) * (const uint8_t*)d;
    case mxINT16_CLASS: return (T) * (const int16_t*)d;
    case mxUINT16_CLASS: return (T) * (const uint16_t*)d;
    case mxINT32_CLASS: return (T) * (const int32_t*)d;
    case mxUINT32_CLASS: return (T) * (const uint32_t*)d;
    case mxINT64_CLASS: return (T) * (const int64_t*)d;
    case mxUINT64_CLASS: return (T) * (const uint64_t*)d;
    case mxSINGLE_CLASS: return (T) * (const float*)d;
    case mxDOUBLE_CLASS: return (T) * (const double*)d;
    default: cx.fail("Unknown input type");
  }
}

template <class T>
std::vector<T> vector_of(const mxArray* in, const Context& cx) {
  if (!in) cx.fail("Null input");
  if (mxIsComplex(in)) cx.fail("Input is complex");
  const size_t n = mxGetNumberOfElements(in);
  std::vector<T> v(n);
  if (mxIsCell(in)) {
    for (size_t i = 0; i < n; ++i) v[i] = scalar<T>(mxGetCell(in, i), cx);
    return v;
  }
  if (n == 0) return v;
  const void* d = mxGetData(in);
  if (!d) cx.fail("Input is invalid");
  auto copy = [&](auto* p) { std::copy(p, p + n, v.begin()); };
  switch (mxGetClassID(in)) {
    case mxLOGICAL_CLASS: copy((const mxLogical*)d); break;
    case mxINT8_CLASS: copy((const int8_t*)d); break;
    case mxUINT8_CLASS: copy((const uint8_t*)d); break;
    case mxINT16_CLASS: copy((const int16_t*)d); break;
    case mxUINT16_CLASS: copy((const uint16_t*)d); break;
    case mxINT32_CLASS: copy((const int32_t*)d); break;
    case mxUINT32_CLASS: copy((const uint32_t*)d); break;
    case mxINT64_CLASS: copy((const int64_t*)d); break;
    case mxUINT64_CLASS: copy((const uint64_t*)d); break;
    case mxSINGLE_CLASS: copy((const float*)d); break;
    case mxDOUBLE_CLASS: copy((const double*)d); break;
    default: cx.fail("Unknown input type");
  }
  return v;
}

CodecBox* codec_of(const mxArray* obj, const Context& cx) {
  if (!obj) cx.fail("Null object");
  const mxArray* h = mxGetProperty(obj, 0, "mexHandle_");
  if (!h || !mxGetData(h) || mxGetClassID(h) != mxUINT64_CLASS || mxGetNumberOfElements(h) != 1)
    cx.fail("Invalid object");
  CodecBox* box = reinterpret_cast<CodecBox*>((uintptr_t) * (const uint64_t*)mxGetData(h));
  if (!box || box->tag != CodecBox::kTag) cx.fail("Null object");
  return box;
}

// wrap/Matlab/util/BitMatrix.h: a sparse matrix's pattern (its stored entries), or the
// nonzeros of a dense double matrix, row by row in ascending columns
fec::SparseBitMatrix matrix_of(const mxArray* in, const Context& cx) {
  if (!in) cx.fail("Null mxArray");
  if (mxIsComplex(in) || !mxGetData(in)) cx.fail("Invalid data");
  const size_t M = mxGetM(in), N = mxGetN(in);
  std::vector<std::vector<int32_t>> rows(M);
  if (mxIsSparse(in)) {
    const mwIndex* jc = mxGetJc(in);
    const mwIndex* ir = mxGetIr(in);
    if (!jc || !ir) cx.fail("Invalid data");
    for (size_t j = 0; j < N; ++j)
      for (mwIndex k = jc[j]; k < jc[j + 1]; ++k) {
        if (ir[k] >= M) cx.fail("Invalid data");
        rows[ir[k]].push_back((int32_t)j);
      }
  } else {
    const double* pr = mxGetPr(in);
    if (!pr) cx.fail("Invalid data");
    for (size_t j = 0; j < N; ++j)
      for (size_t i = 0; i < M; ++i)
        if (pr[j * M + i] != 0) rows[i].push_back((int32_t)j);
  }
  std::vector<int32_t> rp(1, 0), col;
  for (auto& r : rows) {
    col.insert(col.end(), r.begin(), r.end());
    rp.push_back((int32_t)col.size());
  }
  return fec::SparseBitMatrix(std::move(rp), std::move(col), N);
}

// wrap/Matlab/Ldpc/DecoderOptions.h with the map of Conversion.h (struct array {key, value})
fec::Ldpc::DecoderOptions decoder_options_of(const mxArray* in, Context cx) {
  if (!in || !mxIsStruct(in)) cx.fail("Decoder options must be a struct");
  fec::Ldpc::DecoderOptions d;
  d.iterations(scalar<size_t>(mxGetField(in, 0, "iterations"), Context{cx.what + "In iterations: "}));
  d.algorithm((fec::DecoderAlgorithm)scalar<uint32_t>(mxGetField(in, 0, "algorithm"), Context{cx.what + "In algorithm: "}));
  const mxArray* sf = mxGetField(in, 0, "scalingFactor");
  const Context sc{cx.what + "In scaling factor: "};
  if (!sf) sc.fail("Null input");
  std::unordered_map<size_t, std::vector<double>> map;
  for (size_t i = 0; i < mxGetNumberOfElements(sf); ++i)
    map.emplace(scalar<size_t>(mxGetField(sf, i, "key"), sc), vector_of<double>(mxGetField(sf, i, "value"), sc));
  d.scalingFactor(map);
  return d;
}

// ---- C++ -> MATLAB
template <class T>
mxArray* numeric_scalar(mxClassID cls, T v) {
  mxArray* out = mxCreateNumericMatrix(1, 1, cls, mxREAL);
  std::memcpy(mxGetData(out), &v, sizeof(T));
  return out;
}

template <class T, class S>
mxArray* numeric_column(mxClassID cls, const std::vector<S>& v) {
  mxArray* out = mxCreateNumericMatrix(v.size(), 1, cls, mxREAL);
  T* d = (T*)mxGetData(out);
  for (size_t i = 0; i < v.size(); ++i) d[i] = (T)v[i];
  return out;
}

mxArray* handle_array(CodecBox* box) {
  return numeric_scalar<uint64_t>(mxUINT64_CLASS, (uint64_t)(uintptr_t)box);
}

mxArray* sparse_of(const fec::SparseBitMatrix& H) {
  // column-compressed (MATLAB's layout) from the row-compressed matrix
  const size_t M = H.rows(), N = H.cols(), E = H.size();
  mxArray* out = mxCreateSparseLogicalMatrix(M, N, std::max<size_t>(E, 1));
  mwIndex* jc = mxGetJc(out);
  mwIndex* ir = mxGetIr(out);
  mxLogical* pr = (mxLogical*)mxGetData(out);
  std::vector<mwIndex> count(N + 1, 0);
  for (size_t e = 0; e < E; ++e) count[H.at(e) + 1]++;
  for (size_t j = 0; j < N; ++j) count[j + 1] += count[j];
  for (size_t j = 0; j <= N; ++j) jc[j] = count[j];
  const auto& rp = H.rowPtr();
  for (size_t r = 0; r < M; ++r)
    for (int32_t e = rp[r]; e < rp[r + 1]; ++e) ir[count[H.at(e)]++] = r;
  for (size_t e = 0; e < E; ++e) pr[e] = true;
  return out;
}

mxArray* decoder_options_array(const fec::Ldpc::DecoderOptions& d) {
  const char* names[] = {"iterations", "algorithm", "scalingFactor"};
  mxArray* out = mxCreateStructMatrix(1, 1, 3, names);
  mxSetField(out, 0, names[0], numeric_scalar<uint64_t>(mxUINT64_CLASS, (uint64_t)d.iterations()));
  mxSetField(out, 0, names[1], numeric_scalar<uint32_t>(mxUINT32_CLASS, (uint32_t)d.algorithm()));
  const auto map = d.scalingFactor();
  const char* kv[] = {"key", "value"};
  mxArray* sf = mxCreateStructMatrix(1, map.size(), 2, kv);
  size_t i = 0;
  for (const auto& e : map) {
    mxSetField(sf, i, kv[0], numeric_scalar<uint64_t>(mxUINT64_CLASS, (uint64_t)e.first));
    mxSetField(sf, i, kv[1], numeric_column<double>(mxDOUBLE_CLASS, e.second));
    ++i;
  }
  mxSetField(out, 0, names[2], sf);
  return out;
}

const mxArray* arg(int nrhs, const mxArray* prhs[], int i, const char* what) {
  if (i >= nrhs) throw std::invalid_argument(std::string("Missing argument: ") + what);
  return prhs[i];
}

void run(size_t id, int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  auto out = [&](int i, mxArray* v) {
    if (i < std::max(nlhs, 1)) plhs[i] = v;
    else mxDestroyArray(v);
  };
  using Bits = std::vector<fec::BitField<size_t>>;
  switch (id) {
    case Codec_destroy: {
      try {
        CodecBox* box = codec_of(arg(nrhs, prhs, 0, "codec"), Context{});
        box->tag = 0;
        delete box;
        mexUnlock();
      } catch (...) {
      }
      out(0, handle_array(nullptr));
      return;
    }
    case Codec_save: {
      CodecBox* box = codec_of(arg(nrhs, prhs, 0, "codec"), Context{"Save codec"});
      std::vector<char> bytes = box->codec.save();
      // FeCl sizes the buffer from a counting stream that has flushed whole 4096-byte blocks,
      // plus 4096 bytes (src/Serialization.h:67-79); the tail stays zero
      const size_t size = bytes.size() / 4096 * 4096 + 4096;
      mxArray* a = mxCreateNumericMatrix(size, 1, mxUINT8_CLASS, mxREAL);
      std::memcpy(mxGetData(a), bytes.data(), bytes.size());
      out(0, a);
      return;
    }
    case Codec_load: {
      const mxArray* a = arg(nrhs, prhs, 0, "archive");
      if (!mxGetData(a)) throw std::invalid_argument("invalid");
      std::vector<char> bytes = vector_of<char>(a, Context{"Load codec: "});
      std::unique_ptr<CodecBox> box(new CodecBox(fec::Ldpc::load(bytes.data(), bytes.size())));
      mexLock();
      out(0, handle_array(box.release()));
      return;
    }
    case Codec_check: {
      CodecBox* box = codec_of(arg(nrhs, prhs, 0, "codec"), Context{"In check: In Codec object: "});
      Bits parity = vector_of<fec::BitField<size_t>>(arg(nrhs, prhs, 1, "parity"), Context{"In Parity vector: "});
      out(0, mxCreateLogicalScalar(box->codec.check(parity)));
      return;
    }
    case Codec_encode: {
      CodecBox* box = codec_of(arg(nrhs, prhs, 0, "codec"), Context{"In encode: In Codec object: "});
      Bits msg = vector_of<fec::BitField<size_t>>(arg(nrhs, prhs, 1, "msg"), Context{"In Msg vector: "});
      Bits parity;
      box->codec.encode(msg, parity);
      out(0, numeric_column<uint64_t>(mxUINT64_CLASS, parity));
      return;
    }
    case Codec_decode: {
      CodecBox* box = codec_of(arg(nrhs, prhs, 0, "codec"), Context{"In decode: In Codec object: "});
      std::vector<double> parity = vector_of<double>(arg(nrhs, prhs, 1, "parity"), Context{"In Parity vector: "});
      Bits msg;
      box->codec.decode(parity, msg);
      out(0, numeric_column<uint64_t>(mxUINT64_CLASS, msg));
      return;
    }
    case Codec_soDecode: {
      // inputs (parity, [state], [syst]); outputs (msg, [syst], [state], [parity]) as wrap.h:105-135
      CodecBox* box = codec_of(arg(nrhs, prhs, 0, "codec"), Context{"In doDecode: In Codec object: "});
      std::vector<double> parityIn = vector_of<double>(arg(nrhs, prhs, 1, "parity"), Context{"In Parity vector: "});
      std::vector<double> stateIn, systIn;
      if (nrhs > 2) stateIn = vector_of<double>(prhs[2], Context{"In State vector: "});
      if (nrhs > 3) systIn = vector_of<double>(prhs[3], Context{"In Syst vector: "});
      fec::Codec::Input<> input;
      if (!parityIn.empty()) input.parity(parityIn);
      if (!stateIn.empty()) input.state(stateIn);
      if (!systIn.empty()) input.syst(systIn);
      std::vector<double> msg, syst, state, parity;
      fec::Codec::Output<> output;
      output.msg(msg);
      if (nlhs > 1) output.syst(syst);
      if (nlhs > 2) output.state(state);
      if (nlhs > 3) output.parity(parity);
      box->codec.soDecode(input, output);
      out(0, numeric_column<double>(mxDOUBLE_CLASS, msg));
      if (nlhs > 1) out(1, numeric_column<double>(mxDOUBLE_CLASS, syst));
      if (nlhs > 2) out(2, numeric_column<double>(mxDOUBLE_CLASS, state));
      if (nlhs > 3) out(3, numeric_column<double>(mxDOUBLE_CLASS, parity));
      return;
    }
    case Codec_get_msgSize:
      out(0, numeric_scalar<uint64_t>(mxUINT64_CLASS, codec_of(arg(nrhs, prhs, 0, "codec"), Context{})->codec.msgSize()));
      return;
    case Codec_get_systSize:
      out(0, numeric_scalar<uint64_t>(mxUINT64_CLASS, codec_of(arg(nrhs, prhs, 0, "codec"), Context{})->codec.systSize()));
      return;
    case Codec_get_stateSize:
      out(0, numeric_scalar<uint64_t>(mxUINT64_CLASS, codec_of(arg(nrhs, prhs, 0, "codec"), Context{})->codec.stateSize()));
      return;
    case Codec_get_paritySize:
      out(0, numeric_scalar<uint64_t>(mxUINT64_CLASS, codec_of(arg(nrhs, prhs, 0, "codec"), Context{})->codec.paritySize()));
      return;
    case Codec_get_workGroupSize:
      out(0, numeric_scalar<int32_t>(mxINT32_CLASS, codec_of(arg(nrhs, prhs, 0, "codec"), Context{})->codec.getWorkGroupSize()));
      return;
    case Codec_set_workGroupSize:
      codec_of(arg(nrhs, prhs, 0, "codec"), Context{})->codec.setWorkGroupSize(scalar<int>(arg(nrhs, prhs, 1, "size"), Context{}));
      return;
    case Ldpc_constructor: {
      const mxArray* enc = arg(nrhs, prhs, 0, "encoder options");
      if (!mxIsStruct(enc)) throw std::invalid_argument("In Encoder options: Encoder options must be a struct");
      fec::SparseBitMatrix H = matrix_of(mxGetField(enc, 0, "checkMatrix"), Context{"In Encoder options: In check matrix: "});
      fec::Ldpc::DecoderOptions d = decoder_options_of(arg(nrhs, prhs, 1, "decoder options"), Context{"In Decoder options: "});
      std::unique_ptr<CodecBox> box(new CodecBox(fec::Ldpc(fec::Ldpc::EncoderOptions(H), d)));
      mexLock();
      out(0, handle_array(box.release()));
      return;
    }
    case Ldpc_get_decoderOptions:
      out(0, decoder_options_array(codec_of(arg(nrhs, prhs, 0, "codec"), Context{"In Codec object: "})->codec.getDecoderOptions()));
      return;
    case Ldpc_set_decoderOptions:
      codec_of(arg(nrhs, prhs, 0, "codec"), Context{})->codec.setDecoderOptions(
          decoder_options_of(arg(nrhs, prhs, 1, "decoder options"), Context{"In Decoder options: "}));
      return;
    case Ldpc_puncturing: {
      CodecBox* box = codec_of(arg(nrhs, prhs, 0, "codec"), Context{});
      const mxArray* p = arg(nrhs, prhs, 1, "puncture options");
      if (!mxIsStruct(p)) throw std::invalid_argument("In Puncture options: Puncture options must be a struct");
      fec::Ldpc::PunctureOptions po;
      po.mask(vector_of<bool>(mxGetField(p, 0, "mask"), Context{"In Puncture options: In mask: "}));
      po.systMask(vector_of<bool>(mxGetField(p, 0, "systMask"), Context{"In Puncture options: In systematic mask: "}));
      fec::Permutation perm = box->codec.puncturing(po);
      std::vector<uint64_t> idx(perm.outputSize());
      for (size_t i = 0; i < idx.size(); ++i) idx[i] = (uint64_t)perm[i] + 1;
      out(0, numeric_column<uint64_t>(mxUINT64_CLASS, idx));
      return;
    }
    case Ldpc_DvbS2_matrix:
      out(0, sparse_of(fec::Ldpc::DvbS2::matrix(scalar<size_t>(arg(nrhs, prhs, 0, "size"), Context{"In size: "}),
                                                scalar<double>(arg(nrhs, prhs, 1, "rate"), Context{"In rate: "}))));
      return;
    default:
      if (id < kFcnCount) throw std::invalid_argument("Only LDPC codecs are provided by this library");
      throw std::invalid_argument("Unknown fonction id");
  }
}

}  // namespace

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  std::string err, id = "fec:invalid_argument";
  try {
    if (nrhs < 1) throw std::invalid_argument("Missing function id");
    const size_t fcn = scalar<size_t>(prhs[0], Context{"In function id: "});
    run(fcn, nlhs, plhs, nrhs - 1, prhs + 1);
    return;
  } catch (const std::invalid_argument& e) {
    err = e.what();
  } catch (const std::exception& e) {
    err = e.what();
    id = "fec:runtime_error";
  }
  mexErrMsgIdAndTxt(id.c_str(), "%s", err.c_str());
}
