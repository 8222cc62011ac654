// Drives the MATLAB gateway (fecl_amd/matlab/fecl_wrap.cpp) through the MEX test double the way
// FeCl's MATLAB classes call `fec.bin.wrap` (+fec/Codec.m, +fec/Ldpc.m), with FeCl's function
// ids (+fec/+detail/WrapFcnId.m), and compares every result with the C++ API.
//   mex_gateway_test cpu   host-only calls (DvbS2 matrix, errors)
//   mex_gateway_test       everything, on the GPU
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "fecl/Ldpc.hpp"
#include "mex.h"

static int failures = 0;
#define CHECK(c)                                                                        \
  do {                                                                                  \
    if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); ++failures; } \
  } while (0)

enum Id {
  Codec_destroy = 0, Codec_save = 1, Codec_load = 2, Codec_check = 3, Codec_encode = 4, Codec_decode = 5,
  Codec_soDecode = 6, Codec_get_msgSize = 7, Codec_get_systSize = 8, Codec_get_stateSize = 9,
  Codec_get_paritySize = 10, Codec_get_workGroupSize = 11, Codec_set_workGroupSize = 12, Turbo_constructor = 13,
  Ldpc_constructor = 18, Ldpc_get_decoderOptions = 19, Ldpc_set_decoderOptions = 20, Ldpc_puncturing = 21,
  Ldpc_DvbS2_matrix = 22, Trellis_constructor = 27
};

static mxArray* dbl(double v) {
  mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
  *mxGetPr(a) = v;
  return a;
}
template <class T>
static mxArray* column(mxClassID cls, const std::vector<T>& v) {
  mxArray* a = mxCreateNumericMatrix(v.size(), 1, cls, mxREAL);
  if (!v.empty()) std::memcpy(mxGetData(a), v.data(), v.size() * sizeof(T));
  return a;
}
template <class T>
static std::vector<T> values(const mxArray* a) {
  const T* p = (const T*)mxGetData(a);
  return std::vector<T>(p, p + mxGetNumberOfElements(a));
}

// fec.bin.wrap(id, args...) with nout outputs; the arguments are destroyed afterwards
static std::vector<mxArray*> call(int id, std::vector<mxArray*> args, int nout = 1, std::vector<const mxArray*> keep = {}) {
  std::vector<const mxArray*> in;
  mxArray* idarg = mxCreateNumericMatrix(1, 1, mxUINT32_CLASS, mxREAL);
  *(uint32_t*)mxGetData(idarg) = (uint32_t)id;
  in.push_back(idarg);
  for (const mxArray* k : keep) in.push_back(k);
  for (mxArray* a : args) in.push_back(a);
  std::vector<mxArray*> out(std::max(nout, 1), nullptr);
  try {
    mexFunction(nout, out.data(), (int)in.size(), in.data());
  } catch (...) {
    mxDestroyArray(idarg);
    for (mxArray* a : args) mxDestroyArray(a);
    throw;
  }
  mxDestroyArray(idarg);
  for (mxArray* a : args) mxDestroyArray(a);
  return out;
}

static bool throws(int id, std::vector<mxArray*> args, std::vector<const mxArray*> keep = {}, std::string* msg = nullptr) {
  try {
    auto r = call(id, args, 1, keep);
    for (mxArray* a : r) mxDestroyArray(a);
  } catch (const mexsim::Error& e) {
    if (msg) *msg = e.what();
    return e.id.compare(0, 4, "fec:") == 0;
  }
  return false;
}

static mxArray* decoder_struct(double iterations, uint32_t algorithm, double sf) {
  const char* names[] = {"iterations", "algorithm", "scalingFactor"};
  mxArray* d = mxCreateStructMatrix(1, 1, 3, names);
  mxSetField(d, 0, "iterations", dbl(iterations));
  mxArray* alg = mxCreateNumericMatrix(1, 1, mxUINT32_CLASS, mxREAL);
  *(uint32_t*)mxGetData(alg) = algorithm;
  mxSetField(d, 0, "algorithm", alg);
  const char* kv[] = {"key", "value"};
  mxArray* map = mxCreateStructMatrix(1, 1, 2, kv);
  mxSetField(map, 0, "key", dbl(0));
  mxSetField(map, 0, "value", dbl(sf));
  mxSetField(d, 0, "scalingFactor", map);
  return d;
}

static mxArray* encoder_struct(mxArray* H) {
  const char* names[] = {"checkMatrix"};
  mxArray* e = mxCreateStructMatrix(1, 1, 1, names);
  mxSetField(e, 0, "checkMatrix", H);
  return e;
}

// test/Cpp/operations.h:42-59: consistent-Gaussian LLRs around the codeword
static std::vector<double> distort(const std::vector<uint64_t>& cw, double snrdb, unsigned seed) {
  const double snr = std::pow(10.0, snrdb / 10.0);
  std::minstd_rand0 g(seed);
  std::normal_distribution<double> nd(snr * 4.0, 4.0 * std::sqrt(snr / 2.0));
  std::vector<double> llr(cw.size());
  for (size_t i = 0; i < cw.size(); ++i) llr[i] = (cw[i] ? 1.0 : -1.0) * nd(g);
  return llr;
}

static void host_cases() {
  // DvbS2 matrix: sparse logical M x N with the C++ API's pattern
  auto r = call(Ldpc_DvbS2_matrix, {dbl(16200), dbl(0.5)});
  mxArray* H = r[0];
  auto ref = fec::Ldpc::DvbS2::matrix(16200, 0.5);
  CHECK(mxIsSparse(H) && mxGetClassID(H) == mxLOGICAL_CLASS);
  CHECK(mxGetM(H) == ref.rows() && mxGetN(H) == ref.cols());
  const mwIndex* jc = mxGetJc(H);
  const mwIndex* ir = mxGetIr(H);
  CHECK(jc[ref.cols()] == ref.size());
  std::vector<std::vector<size_t>> rows(ref.rows());
  for (size_t j = 0; j < ref.cols(); ++j)
    for (mwIndex k = jc[j]; k < jc[j + 1]; ++k) rows[ir[k]].push_back(j);
  bool same = true;
  for (size_t rr = 0; rr < ref.rows(); ++rr) {
    const auto& rp = ref.rowPtr();
    std::vector<size_t> want(ref.colIdx().begin() + rp[rr], ref.colIdx().begin() + rp[rr + 1]);
    same &= rows[rr] == want;
  }
  CHECK(same);
  mxDestroyArray(H);
  // errors come back through mexErrMsgIdAndTxt
  std::string msg;
  CHECK(throws(Ldpc_DvbS2_matrix, {dbl(1000), dbl(0.5)}, {}, &msg) && msg.find("Invalid length") != std::string::npos);
  CHECK(throws(Ldpc_DvbS2_matrix, {dbl(64800), dbl(0.7)}, {}, &msg) && msg.find("Invalid rate") != std::string::npos);
  CHECK(throws(Turbo_constructor, {dbl(1), dbl(2)}, {}, &msg) && msg.find("Only LDPC") != std::string::npos);
  CHECK(throws(Trellis_constructor, {}, {}));
  CHECK(throws(99, {}));
  CHECK(throws(Ldpc_DvbS2_matrix, {dbl(64800)}));   // missing argument
  mxArray* bad_handle = column<uint64_t>(mxUINT64_CLASS, {0});
  mxArray* obj = mexsim::object_with_property("mexHandle_", bad_handle);
  CHECK(throws(Codec_get_msgSize, {}, {obj}, &msg) && msg.find("Null object") != std::string::npos);
  mxDestroyArray(obj);
}

static void device_cases() {
  auto h = call(Ldpc_DvbS2_matrix, {dbl(16200), dbl(0.5)});
  auto ctor = call(Ldpc_constructor, {encoder_struct(h[0]), decoder_struct(30, 2, 1.0)});
  CHECK(mxGetClassID(ctor[0]) == mxUINT64_CLASS);
  mxArray* obj = mexsim::object_with_property("mexHandle_", ctor[0]);
  fec::Ldpc ref(fec::Ldpc::EncoderOptions(fec::Ldpc::DvbS2::matrix(16200, 0.5)),
                fec::Ldpc::DecoderOptions().algorithm(fec::Approximate).iterations(30));
  auto size_of = [&](int id) {
    auto r = call(id, {}, 1, {obj});
    const uint64_t v = *(const uint64_t*)mxGetData(r[0]);
    mxDestroyArray(r[0]);
    return v;
  };
  CHECK(size_of(Codec_get_msgSize) == ref.msgSize() && size_of(Codec_get_systSize) == ref.systSize());
  CHECK(size_of(Codec_get_paritySize) == ref.paritySize() && size_of(Codec_get_stateSize) == ref.stateSize());
  // encode / check / decode (messages as MATLAB doubles, bits back as uint64)
  const size_t K = ref.msgSize(), N = ref.paritySize(), F = 3;
  std::mt19937 g(5);
  std::vector<double> msg(K * F);
  for (auto& b : msg) b = (double)(g() & 1);
  auto enc = call(Codec_encode, {column<double>(mxDOUBLE_CLASS, msg)}, 1, {obj});
  CHECK(mxGetClassID(enc[0]) == mxUINT64_CLASS && mxGetNumberOfElements(enc[0]) == N * F);
  std::vector<uint64_t> cw = values<uint64_t>(enc[0]);
  std::vector<fec::BitField<size_t>> msgb(msg.begin(), msg.end()), cwref;
  ref.encode(msgb, cwref);
  CHECK(std::equal(cw.begin(), cw.end(), cwref.begin(), [](uint64_t a, fec::BitField<size_t> b) { return a == (size_t)b; }));
  auto chk = call(Codec_check, {column<uint64_t>(mxUINT64_CLASS, cw)}, 1, {obj});
  CHECK(mxGetClassID(chk[0]) == mxLOGICAL_CLASS && *(mxLogical*)mxGetData(chk[0]));
  mxDestroyArray(chk[0]);
  mxDestroyArray(enc[0]);
  std::vector<double> llr = distort(cw, 2.0, 1);
  auto dec = call(Codec_decode, {column<double>(mxDOUBLE_CLASS, llr)}, 1, {obj});
  std::vector<uint64_t> bits = values<uint64_t>(dec[0]);
  CHECK(bits.size() == K * F);
  bool ok = true;
  for (size_t i = 0; i < bits.size(); ++i) ok &= bits[i] == (uint64_t)msg[i];
  CHECK(ok);
  mxDestroyArray(dec[0]);
  // soDecode: (parity, state, syst) in; (msg, syst, state, parity) out, as FeCl's wrapper orders them
  std::vector<double> low = distort(cw, -3.0, 2);
  auto so = call(Codec_soDecode, {column<double>(mxDOUBLE_CLASS, low)}, 4, {obj});
  std::vector<double> m, sy, st, pa;
  ref.soDecode(fec::Codec::Input<>().parity(low), fec::Codec::Output<>().msg(m).syst(sy).state(st).parity(pa));
  CHECK(values<double>(so[0]) == m && values<double>(so[1]) == sy);
  CHECK(values<double>(so[2]) == st && values<double>(so[3]) == pa);
  auto so2 = call(Codec_soDecode, {column<double>(mxDOUBLE_CLASS, low), column<double>(mxDOUBLE_CLASS, st),
                                   column<double>(mxDOUBLE_CLASS, sy)}, 1, {obj});
  std::vector<double> m2;
  ref.soDecode(fec::Codec::Input<>().parity(low).state(st).syst(sy), fec::Codec::Output<>().msg(m2));
  CHECK(values<double>(so2[0]) == m2);
  for (mxArray* a : so) mxDestroyArray(a);
  mxDestroyArray(so2[0]);
  // decoder options round trip, work group size
  auto opt = call(Ldpc_get_decoderOptions, {}, 1, {obj});
  CHECK(*(uint64_t*)mxGetData(mxGetField(opt[0], 0, "iterations")) == 30);
  CHECK(*(uint32_t*)mxGetData(mxGetField(opt[0], 0, "algorithm")) == 2);
  mxDestroyArray(opt[0]);
  call(Ldpc_set_decoderOptions, {decoder_struct(12, 1, 0.75)}, 0, {obj});
  auto opt2 = call(Ldpc_get_decoderOptions, {}, 1, {obj});
  CHECK(*(uint64_t*)mxGetData(mxGetField(opt2[0], 0, "iterations")) == 12);
  CHECK(*(uint32_t*)mxGetData(mxGetField(opt2[0], 0, "algorithm")) == 1);
  const mxArray* sf = mxGetField(opt2[0], 0, "scalingFactor");
  CHECK(mxGetNumberOfElements(sf) == 1 && *mxGetPr(mxGetField(sf, 0, "value")) == 0.75);
  mxDestroyArray(opt2[0]);
  call(Codec_set_workGroupSize, {dbl(4)}, 0, {obj});
  auto wgs = call(Codec_get_workGroupSize, {}, 1, {obj});
  CHECK(mxGetClassID(wgs[0]) == mxINT32_CLASS && *(int32_t*)mxGetData(wgs[0]) == 4);
  mxDestroyArray(wgs[0]);
  // puncturing: 1-based indices of FeCl's sequence
  const char* pn[] = {"mask", "systMask"};
  mxArray* po = mxCreateStructMatrix(1, 1, 2, pn);
  mxSetField(po, 0, "mask", column<double>(mxDOUBLE_CLASS, {1, 0}));
  mxSetField(po, 0, "systMask", column<double>(mxDOUBLE_CLASS, {1}));
  auto perm = call(Ldpc_puncturing, {po}, 1, {obj});
  auto pref = ref.puncturing(fec::Ldpc::PunctureOptions().mask({1, 0}).systMask({1}));
  std::vector<uint64_t> idx = values<uint64_t>(perm[0]);
  ok = idx.size() == pref.outputSize();
  for (size_t i = 0; ok && i < idx.size(); ++i) ok &= idx[i] == pref[i] + 1;
  CHECK(ok);
  mxDestroyArray(perm[0]);
  // save / load: FeCl's archive bytes, padded as FeCl pads them; the loaded codec decodes alike
  auto saved = call(Codec_save, {}, 1, {obj});
  fec::Ldpc tuned(fec::Ldpc::EncoderOptions(fec::Ldpc::DvbS2::matrix(16200, 0.5)),
                  fec::Ldpc::DecoderOptions().algorithm(fec::Linear).iterations(12).scalingFactor(0.75), 4);
  std::vector<char> want = tuned.save();
  CHECK(mxGetClassID(saved[0]) == mxUINT8_CLASS && mxGetNumberOfElements(saved[0]) == want.size() / 4096 * 4096 + 4096);
  CHECK(std::memcmp(mxGetData(saved[0]), want.data(), want.size()) == 0);
  auto loaded = call(Codec_load, {saved[0]});
  mxArray* obj2 = mexsim::object_with_property("mexHandle_", loaded[0]);
  auto d1 = call(Codec_decode, {column<double>(mxDOUBLE_CLASS, low)}, 1, {obj});
  auto d2 = call(Codec_decode, {column<double>(mxDOUBLE_CLASS, low)}, 1, {obj2});
  CHECK(values<uint64_t>(d1[0]) == values<uint64_t>(d2[0]));
  mxDestroyArray(d1[0]);
  mxDestroyArray(d2[0]);
  // a dense double check matrix with dependent checks (Gallager (30,3,5)): FeCl's form, K = 14
  auto G = fec::Ldpc::Gallager::matrix(30, 3, 5, 0);
  mxArray* dense = mxCreateDoubleMatrix(G.rows(), G.cols(), mxREAL);
  for (size_t rr = 0; rr < G.rows(); ++rr)
    for (int32_t e = G.rowPtr()[rr]; e < G.rowPtr()[rr + 1]; ++e) mxGetPr(dense)[G.at(e) * G.rows() + rr] = 1.0;
  auto gc = call(Ldpc_constructor, {encoder_struct(dense), decoder_struct(10, 2, 1.0)});
  mxArray* obj3 = mexsim::object_with_property("mexHandle_", gc[0]);
  CHECK(size_of(Codec_get_msgSize) == K);
  auto gk = call(Codec_get_msgSize, {}, 1, {obj3});
  CHECK(*(uint64_t*)mxGetData(gk[0]) == 14);
  mxDestroyArray(gk[0]);
  // destroy returns a null handle; the object then no longer names a codec
  for (mxArray* o : {obj, obj2, obj3}) {
    auto z = call(Codec_destroy, {}, 1, {o});
    CHECK(*(uint64_t*)mxGetData(z[0]) == 0);
    mxDestroyArray(z[0]);
    mxDestroyArray(o);
  }
}

int main(int argc, char** argv) {
  const long live0 = mexsim::live_arrays();
  try {
    host_cases();
    if (!(argc > 1 && std::string(argv[1]) == "cpu")) device_cases();
  } catch (const std::exception& e) {
    std::printf("FAIL: unexpected exception: %s\n", e.what());
    ++failures;
  }
  CHECK(mexsim::live_arrays() == live0);
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
  return failures ? 1 : 0;
}
