// C++ facade test: FeCl's own LDPC test cases (test/Cpp/Ldpc.cpp:51-113, test/Cpp/operations.h)
// written against include/fecl/Ldpc.hpp, on FeCl's test code Gallager(2048, 3, 5, seed 0) and on
// a DVB-S2 code.  Needs a GPU.
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../include/fecl/Ldpc.hpp"

static int failures = 0;
#define CHECK(c)                                                      \
  do {                                                                \
    if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); ++failures; } \
  } while (0)

// test/Cpp/operations.h:42-59: consistent-Gaussian LLRs around the codeword, seed 0
static std::vector<double> distort(const std::vector<fec::BitField<size_t>>& in, double snrdb) {
  const int bpsk[2] = {-1, 1};
  double snr = std::pow(10.0, snrdb / 10.0);
  std::minstd_rand0 g(0);
  std::normal_distribution<double> nd(snr * 4.0, 4.0 * std::sqrt(snr / 2.0));
  std::vector<double> llr(in.size());
  for (size_t i = 0; i < in.size(); ++i) llr[i] = bpsk[(size_t)in[i]] * nd(g);
  return llr;
}

template <class F>
static void expect_throw(F f, const char* what) {
  try { f(); } catch (std::invalid_argument&) { return; }
  std::printf("FAIL: no std::invalid_argument for %s\n", what);
  ++failures;
}

static void reference_cases(const fec::SparseBitMatrix& H) {
  for (auto alg : {fec::Exact, fec::Linear, fec::Approximate}) {
    auto dec = fec::Ldpc::DecoderOptions().algorithm(alg).iterations(30);
    fec::Ldpc code(fec::Ldpc::EncoderOptions(H), dec);
    // encode / check (test_encode)
    std::vector<fec::BitField<size_t>> msg(code.msgSize() * 5, 1), parity;
    code.encode(msg, parity);
    CHECK(code.check(parity));
    // decode recovers the all-ones message at 2 dB (test_decode, n = 5)
    auto llr = distort(parity, 2.0);
    std::vector<fec::BitField<size_t>> out;
    code.decode(llr, out);
    CHECK(out.size() == msg.size());
    size_t errs = 0;
    for (size_t i = 0; i < out.size(); ++i) errs += (size_t)out[i] != (size_t)msg[i];
    CHECK(errs == 0);
    // soDecode hard decisions equal decode (test_soDecode)
    std::vector<double> m;
    code.soDecode(fec::Codec::Input<>().parity(llr), fec::Codec::Output<>().msg(m));
    size_t diff = 0;
    for (size_t i = 0; i < m.size(); ++i) diff += (m[i] >= 0) != ((size_t)out[i] != 0);
    CHECK(diff == 0);
    // syst extrinsic == parity extrinsic prefix (test_ldpc_soDecode_systOut, -5 dB)
    auto llr5 = distort(parity, -5.0);
    std::vector<double> po, so;
    code.soDecode(fec::Codec::Input<>().parity(llr5), fec::Codec::Output<>().parity(po).syst(so));
    bool same = true;
    for (size_t f = 0; f < 5; ++f)
      for (size_t i = 0; i < code.systSize(); ++i)
        same &= so[f * code.systSize() + i] == po[f * code.paritySize() + i];
    CHECK(same);
    // zero a-priori state / syst inputs change nothing (test_soDecode_0stateIn / 0systIn)
    std::vector<double> zs(code.stateSize() * 5, 0.0), zy(code.systSize() * 5, 0.0), m0, m1;
    code.soDecode(fec::Codec::Input<>().parity(llr5), fec::Codec::Output<>().msg(m0));
    code.soDecode(fec::Codec::Input<>().parity(llr5).syst(zy), fec::Codec::Output<>().msg(m1));
    CHECK(m0 == m1);
    // size errors (test_decode_badParitySize, test_soDecode_bad*)
    std::vector<double> bad(code.paritySize() + 1);
    expect_throw([&] { code.decode(bad); }, "decode bad parity");
    std::vector<double> badsyst(3), st;
    expect_throw([&] { code.soDecode(fec::Codec::Input<>().parity(llr5).syst(badsyst), fec::Codec::Output<>().msg(st)); },
                 "bad syst");
    expect_throw([&] { code.soDecode(fec::Codec::Input<>().parity(llr5).state(badsyst), fec::Codec::Output<>().msg(st)); },
                 "bad state");
    std::vector<double> empty;
    expect_throw([&] { code.soDecode(fec::Codec::Input<>().syst(empty), fec::Codec::Output<>().msg(st)); }, "no parity");
    std::vector<fec::BitField<size_t>> badmsg(code.msgSize() + 1);
    expect_throw([&] { code.encode(badmsg); }, "bad message");
  }
}

int main() {
  auto G = fec::Ldpc::Gallager::matrix(2048, 3, 5, 0);
  CHECK(G.rows() == 2048 / 5 * 3 && G.cols() == 2048 && G.size() == G.rows() * 5);
  reference_cases(G);
  auto H = fec::Ldpc::DvbS2::matrix(16200, 0.5);
  reference_cases(H);
  {
    // puncturing (test_decode_puncture, test/Cpp/operations.h:130-145): transmit the permuted
    // codeword, dePermute the received LLRs (punctured positions enter as 0), decode
    fec::Ldpc code(fec::Ldpc::EncoderOptions(H), fec::Ldpc::DecoderOptions().algorithm(fec::Approximate).iterations(50));
    auto perm = code.puncturing(fec::Ldpc::PunctureOptions().mask({1, 0}).systMask({1}));
    CHECK(perm.inputSize() == code.paritySize() && perm.outputSize() < code.paritySize());
    std::vector<fec::BitField<size_t>> msg(code.msgSize() * 3, 1), parity;
    code.encode(msg, parity);
    auto tx = perm.permute(parity);
    CHECK(tx.size() == perm.outputSize() * 3);
    auto rx = distort(tx, 4.0);
    auto llr = perm.dePermute(rx);
    CHECK(llr.size() == code.paritySize() * 3);
    size_t zeros = 0;
    for (double x : llr) zeros += x == 0.0;
    CHECK(zeros == (code.paritySize() - perm.outputSize()) * 3);
    std::vector<fec::BitField<size_t>> out;
    code.decode(llr, out);
    size_t errs = 0;
    for (size_t i = 0; i < out.size(); ++i) errs += (size_t)out[i] != (size_t)msg[i];
    CHECK(errs == 0);
  }
  {
    // a code with dependent checks carries more message bits than N - M (src/detail/Ldpc.cpp:307)
    fec::Ldpc code(fec::Ldpc::EncoderOptions(fec::Ldpc::Gallager::matrix(1000, 3, 5, 1)),
                   fec::Ldpc::DecoderOptions().iterations(30));
    CHECK(code.msgSize() == 402);
    std::mt19937 g(3);
    std::vector<fec::BitField<size_t>> msg(code.msgSize() * 4), parity, out;
    for (auto& b : msg) b = g() & 1;
    code.encode(msg, parity);
    CHECK(code.check(parity));
    code.decode(distort(parity, 4.0), out);
    CHECK(out == msg);
  }
  {
    // save / load (test_saveLoad, test/Cpp/operations.h:106-111): the loaded codec decodes like
    // the saved one and saves the same bytes
    fec::Ldpc code(fec::Ldpc::EncoderOptions(G), fec::Ldpc::DecoderOptions().algorithm(fec::Linear).iterations(20), 4);
    std::vector<char> archive(fec::detail::archiveSize(&code) + 4096);
    fec::detail::save(&code, archive.data(), archive.size());
    auto recover = fec::detail::load<fec::Ldpc>(archive.data(), archive.size());
    CHECK(recover->getWorkGroupSize() == 4 && recover->msgSize() == code.msgSize());
    CHECK(recover->getDecoderOptions().algorithm() == fec::Linear && recover->getDecoderOptions().iterations() == 20);
    CHECK(recover->save() == code.save());
    std::vector<fec::BitField<size_t>> msg(code.msgSize() * 2, 1), parity;
    code.encode(msg, parity);
    std::vector<fec::BitField<size_t>> parity2;
    recover->encode(msg, parity2);
    CHECK(parity == parity2);
    auto llr = distort(parity, 1.0);
    std::vector<double> a, b;
    code.soDecode(fec::Codec::Input<>().parity(llr), fec::Codec::Output<>().msg(a));
    recover->soDecode(fec::Codec::Input<>().parity(llr), fec::Codec::Output<>().msg(b));
    CHECK(a == b);
    std::vector<char> junk(100, 7);
    expect_throw([&] { fec::detail::load<fec::Ldpc>(junk.data(), junk.size()); }, "bad archive");
  }
  expect_throw([] { fec::Ldpc::Gallager::matrix(10, 3, 11); }, "bad Gallager shape");
  expect_throw([] { fec::Ldpc::DvbS2::matrix(1000, 0.5); }, "bad length");
  expect_throw([] { fec::Ldpc::DvbS2::matrix(64800, 0.7); }, "bad rate");
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
  return failures ? 1 : 0;
}
