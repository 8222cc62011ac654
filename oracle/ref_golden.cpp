// Golden-vector driver linked against stock FeCl (built by oracle/ref.mk into oracle/_ref/).
// It is test infrastructure only: it writes the input/output vectors that pin the CPU
// restatement in oracle/ and, through it, the HIP decoder.  Raw little-endian files:
//   CSR  : int64 M, N, K, E, row_ptr[M+1], col[E]
//   LLR  : int64 F, N, double[F*N]
//   outs : int64 F, L, payload[F*L]   (double, or uint8 for hard bits, int32 for counts)
//
// Modes
//   csr-dvbs2    <n> <rate> <out>                  DvbS2::matrix(n, rate) after Structure()
//   csr-gallager <n> <wc> <wr> <seed> <out>        Structure(Gallager::matrix(...)).checks()
//   decode <code> <llr> <alg> <iters> <sf> <prec> <outprefix> [state_in] [syst_in]
//          code = dvbs2:<n>:<rate> | gallager:<n>:<wc>:<wr>:<seed>
//          alg  = 0 Exact, 1 Linear, 2 Approximate;  prec = 64 (stock) or 32 (FeCl's
//          BpDecoderImpl instantiated with a float LLR type)
//          sf   = a single factor, or v:<f0>,<f1>,... (one per iteration)
//          writes <outprefix>.hard (uint8 [F][K]), .msg/.parity/.state (double)
//   iters  <code> <llr> <alg> <imax> <sf> <prec> <out>
//          per-frame CN-update count, derived from where soDecode's state output stops
//          changing as the iteration budget grows (FeCl does not expose the count).
//   gallager-raw <n> <wc> <wr> <seed> <out>       Gallager::matrix(...) as built (CSR, K = N-M)
//   structure <csr> <out>                          Structure(EncoderOptions(H)).checks() and
//          msgSize() for the H read from <csr> (CSR layout above; rows listed in order)
//   encode <csr> <msg> <out>                       Structure::encode of every frame: msg file
//          is int64 F, K, uint8[F*K]; writes int64 F, N, uint8[F*N]
//   save <csr> <alg> <iters> <sf> <wgs> <out>     fec::detail::save of fec::Ldpc(Structure(H, options),
//          wgs) as a fec::Codec*, types registered as <Convolutional, Turbo, Ldpc> (the Matlab
//          wrapper's registration, wrap/wrap.h:35); writes the archive bytes only
//   load <archive> <out>                           fec::detail::load<fec::Codec> of those bytes;
//          writes int64 msg, syst, parity, state sizes, workGroupSize, algorithm, iterations,
//          then checks() as CSR (layout above, K = msgSize) appended
//   puncture <code> <mask> <systmask> <out>
//          Structure::puncturing(PunctureOptions().mask(..).systMask(..)): masks are comma
//          lists of 0/1 ("-" = empty); writes int64 inputSize, outputSize, sequence[]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <string>
#include <vector>
#include <memory>
#include <stdexcept>
#include <sstream>

#include "Serialization.h"   // first: archives use typeid type info for the standard containers
#include "Ldpc.h"
#include "Turbo.h"
#include "Convolutional.h"
#include "detail/LlrMetrics.h"
#include "BpDecoderImpl.cpp"   // FeCl's template definitions, instantiated below for float

namespace fec { namespace detail {
struct Fp32LlrMetrics {
  using Type = float;
  static inline Type max() { return std::numeric_limits<Type>::infinity(); }
};
template class BpDecoderImpl<Fp32LlrMetrics, BoxSum>;
template class BpDecoderImpl<Fp32LlrMetrics, LinearBoxSum>;
template class BpDecoderImpl<Fp32LlrMetrics, MinBoxSum>;
}}

using fec::SparseBitMatrix;
namespace fd = fec::detail;

static std::vector<std::string> split(const std::string& s, char c) {
  std::vector<std::string> out; std::string cur;
  for (char ch : s) { if (ch == c) { out.push_back(cur); cur.clear(); } else cur += ch; }
  out.push_back(cur); return out;
}

static SparseBitMatrix make_code(const std::string& spec) {
  auto p = split(spec, ':');
  if (p[0] == "dvbs2") return fec::Ldpc::DvbS2::matrix(std::stoul(p[1]), std::stod(p[2]));
  if (p[0] == "gallager")
    return fec::Ldpc::Gallager::matrix(std::stoul(p[1]), std::stoul(p[2]), std::stoul(p[3]), std::stoull(p[4]));
  throw std::invalid_argument("bad code spec");
}

static void write_csr(const SparseBitMatrix& H, size_t K, const char* path) {
  FILE* f = fopen(path, "wb");
  int64_t hdr[4] = {(int64_t)H.rows(), (int64_t)H.cols(), (int64_t)K, (int64_t)H.size()};
  fwrite(hdr, 8, 4, f);
  int64_t acc = 0;
  fwrite(&acc, 8, 1, f);
  for (auto row = H.begin(); row < H.end(); ++row) { acc += row->size(); fwrite(&acc, 8, 1, f); }
  for (size_t e = 0; e < H.size(); ++e) { int64_t c = (int64_t)H.at(e); fwrite(&c, 8, 1, f); }
  fclose(f);
}

// CSR writer that walks the rows in order (a matrix whose rows were swapped keeps its edges
// where they were stored, so the storage order is not used here).
static void write_csr_rows(const SparseBitMatrix& H, size_t K, const char* path) {
  FILE* f = fopen(path, "wb");
  int64_t hdr[4] = {(int64_t)H.rows(), (int64_t)H.cols(), (int64_t)K, (int64_t)H.size()};
  fwrite(hdr, 8, 4, f);
  int64_t acc = 0;
  fwrite(&acc, 8, 1, f);
  for (auto row = H.begin(); row < H.end(); ++row) { acc += row->size(); fwrite(&acc, 8, 1, f); }
  for (auto row = H.begin(); row < H.end(); ++row)
    for (auto e = row->begin(); e < row->end(); ++e) { int64_t c = (int64_t)*e; fwrite(&c, 8, 1, f); }
  fclose(f);
}

static SparseBitMatrix read_csr(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) throw std::runtime_error("cannot open csr file");
  int64_t hdr[4];
  if (fread(hdr, 8, 4, f) != 4) throw std::runtime_error("short csr file");
  std::vector<int64_t> rp(hdr[0] + 1), col(hdr[3]);
  if (fread(rp.data(), 8, rp.size(), f) != rp.size() || fread(col.data(), 8, col.size(), f) != col.size())
    throw std::runtime_error("short csr file");
  fclose(f);
  std::vector<size_t> sizes(hdr[0]);
  for (int64_t r = 0; r < hdr[0]; ++r) sizes[r] = (size_t)(rp[r + 1] - rp[r]);
  SparseBitMatrix H(sizes, (size_t)hdr[1]);
  auto row = H.begin();
  for (int64_t r = 0; r < hdr[0]; ++r, ++row)
    for (int64_t e = rp[r]; e < rp[r + 1]; ++e) row->set((size_t)col[e]);
  return H;
}

static std::vector<double> read_llr(const char* path, int64_t& F, int64_t& N) {
  FILE* f = fopen(path, "rb");
  if (!f) throw std::runtime_error("cannot open llr file");
  int64_t hdr[2];
  if (fread(hdr, 8, 2, f) != 2) throw std::runtime_error("short llr file");
  F = hdr[0]; N = hdr[1];
  std::vector<double> v(F * N);
  if (fread(v.data(), 8, v.size(), f) != v.size()) throw std::runtime_error("short llr file");
  fclose(f);
  return v;
}

template <class T>
static void write_out(const std::string& path, int64_t F, int64_t L, const T* data) {
  FILE* f = fopen(path.c_str(), "wb");
  int64_t hdr[2] = {F, L};
  fwrite(hdr, 8, 2, f);
  fwrite(data, sizeof(T), (size_t)(F * L), f);
  fclose(f);
}

static fec::Ldpc::DecoderOptions make_options(int alg, size_t iters, const std::string& sf) {
  auto opt = fec::Ldpc::DecoderOptions().algorithm(fec::DecoderAlgorithm(alg)).iterations(iters);
  if (sf.compare(0, 2, "v:") == 0) {
    std::vector<double> v;
    for (auto& s : split(sf.substr(2), ',')) v.push_back(std::stod(s));
    opt.scalingFactor(std::unordered_map<size_t, std::vector<double>>{{0, v}});
  } else {
    opt.scalingFactor(std::stod(sf));
  }
  return opt;
}

static std::unique_ptr<fd::BpDecoder> make_decoder(const fd::Ldpc::Structure& s, int alg, int prec) {
  if (prec == 64) return fd::BpDecoder::create(s);
  switch (alg) {
    case 0: return std::unique_ptr<fd::BpDecoder>(new fd::BpDecoderImpl<fd::Fp32LlrMetrics, fd::BoxSum>(s));
    case 1: return std::unique_ptr<fd::BpDecoder>(new fd::BpDecoderImpl<fd::Fp32LlrMetrics, fd::LinearBoxSum>(s));
    default: return std::unique_ptr<fd::BpDecoder>(new fd::BpDecoderImpl<fd::Fp32LlrMetrics, fd::MinBoxSum>(s));
  }
}

struct SoOut { std::vector<double> msg, parity, state; };

static SoOut so_decode(const fd::Ldpc::Structure& s, int alg, int prec, const std::vector<double>& llr,
                       int64_t F, const std::vector<double>* state_in, const std::vector<double>* syst_in) {
  auto dec = make_decoder(s, alg, prec);
  SoOut o;
  o.msg.resize(F * s.msgSize());
  o.parity.resize(F * s.paritySize());
  o.state.resize(F * s.stateSize());
  fd::Codec::InputIterator in(&s);
  in.parity(llr.begin());
  if (state_in) in.state(state_in->begin());
  if (syst_in) in.syst(syst_in->begin());
  fd::Codec::OutputIterator out(&s);
  out.msg(o.msg.begin()).parity(o.parity.begin()).state(o.state.begin());
  dec->soDecodeBlocks(in, out, F);
  return o;
}

int main(int argc, char** argv) {
  try {
    if (argc < 2) { fprintf(stderr, "usage: see header\n"); return 2; }
    std::string mode = argv[1];
    if (mode == "csr-dvbs2" && argc == 5) {
      auto H = fec::Ldpc::DvbS2::matrix(std::stoul(argv[2]), std::stod(argv[3]));
      fd::Ldpc::Structure s{fec::Ldpc::EncoderOptions(H), make_options(2, 1, "1")};
      write_csr(H, H.cols() - H.rows(), (std::string(argv[4]) + ".raw").c_str());
      write_csr(s.checks(), s.msgSize(), argv[4]);
      return 0;
    }
    if (mode == "csr-gallager" && argc == 7) {
      auto H = fec::Ldpc::Gallager::matrix(std::stoul(argv[2]), std::stoul(argv[3]), std::stoul(argv[4]), std::stoull(argv[5]));
      fd::Ldpc::Structure s{fec::Ldpc::EncoderOptions(H), make_options(2, 1, "1")};
      write_csr(s.checks(), s.msgSize(), argv[6]);
      return 0;
    }
    if ((mode == "decode" && argc >= 9) || (mode == "iters" && argc == 9)) {
      auto H = make_code(argv[2]);
      int64_t F, N;
      auto llr = read_llr(argv[3], F, N);
      int alg = atoi(argv[4]);
      size_t iters = std::stoul(argv[5]);
      std::string sf = argv[6];
      int prec = atoi(argv[7]);
      fd::Ldpc::Structure s{fec::Ldpc::EncoderOptions(H), make_options(alg, iters, sf)};
      if ((int64_t)s.paritySize() != N) throw std::invalid_argument("llr length != code length");
      if (mode == "decode") {
        std::vector<double> state_in, syst_in;
        int64_t f2, l2;
        if (argc >= 10 && std::string(argv[9]) != "-") state_in = read_llr(argv[9], f2, l2);
        if (argc >= 11 && std::string(argv[10]) != "-") syst_in = read_llr(argv[10], f2, l2);
        SoOut o = so_decode(s, alg, prec, llr, F, state_in.empty() ? nullptr : &state_in,
                            syst_in.empty() ? nullptr : &syst_in);
        std::string pre = argv[8];
        write_out(pre + ".msg", F, s.msgSize(), o.msg.data());
        write_out(pre + ".parity", F, s.paritySize(), o.parity.data());
        write_out(pre + ".state", F, s.stateSize(), o.state.data());
        // Hard decisions through FeCl's decodeBlock (its own output path).
        auto dec = make_decoder(s, alg, prec);
        std::vector<fec::BitField<size_t>> hard(F * s.msgSize());
        dec->decodeBlocks(llr.begin(), hard.begin(), F);
        std::vector<uint8_t> h8(hard.size());
        for (size_t i = 0; i < hard.size(); ++i) h8[i] = hard[i] ? 1 : 0;
        write_out(pre + ".hard", F, s.msgSize(), h8.data());
        return 0;
      }
      // iters: smallest I0 with state(I) == state(imax) for every I in [I0, imax].
      std::vector<std::vector<double>> states(iters + 1);
      for (size_t I = 1; I <= iters; ++I) {
        s.setDecoderOptions(make_options(alg, I, sf));
        states[I] = so_decode(s, alg, prec, llr, F, nullptr, nullptr).state;
      }
      std::vector<int32_t> count(F);
      size_t E = s.stateSize();
      for (int64_t f = 0; f < F; ++f) {
        size_t I0 = iters;
        for (size_t I = iters; I >= 1; --I) {
          if (memcmp(&states[I][f * E], &states[iters][f * E], E * 8) != 0) break;
          I0 = I;
        }
        count[f] = (int32_t)I0;
      }
      write_out(std::string(argv[8]), F, 1, count.data());
      return 0;
    }
    if (mode == "gallager-raw" && argc == 7) {
      auto H = fec::Ldpc::Gallager::matrix(std::stoul(argv[2]), std::stoul(argv[3]), std::stoul(argv[4]), std::stoull(argv[5]));
      write_csr_rows(H, H.cols() - H.rows(), argv[6]);
      return 0;
    }
    if (mode == "structure" && argc == 4) {
      fd::Ldpc::Structure s{fec::Ldpc::EncoderOptions(read_csr(argv[2])), make_options(2, 1, "1")};
      write_csr_rows(s.checks(), s.msgSize(), argv[3]);
      return 0;
    }
    if (mode == "encode" && argc == 5) {
      fd::Ldpc::Structure s{fec::Ldpc::EncoderOptions(read_csr(argv[2])), make_options(2, 1, "1")};
      FILE* f = fopen(argv[3], "rb");
      if (!f) throw std::runtime_error("cannot open msg file");
      int64_t hdr[2];
      if (fread(hdr, 8, 2, f) != 2 || hdr[1] != (int64_t)s.msgSize()) throw std::runtime_error("msg file does not fit the code");
      std::vector<uint8_t> m(hdr[0] * hdr[1]);
      if (fread(m.data(), 1, m.size(), f) != m.size()) throw std::runtime_error("short msg file");
      fclose(f);
      std::vector<fec::BitField<size_t>> msg(m.begin(), m.end()), cw(hdr[0] * s.paritySize());
      for (int64_t i = 0; i < hdr[0]; ++i)
        s.encode(msg.begin() + i * s.msgSize(), cw.begin() + i * s.paritySize());
      std::vector<uint8_t> out(cw.size());
      for (size_t i = 0; i < cw.size(); ++i) out[i] = (uint8_t)(size_t)cw[i];
      write_out(argv[4], hdr[0], (int64_t)s.paritySize(), out.data());
      return 0;
    }
    if (mode == "save" && argc == 8) {
      fd::Ldpc::Structure s{fec::Ldpc::EncoderOptions(read_csr(argv[2])),
                            make_options(atoi(argv[3]), std::stoul(argv[4]), argv[5])};
      fec::Ldpc codec(s, atoi(argv[6]));
      const fec::Codec* base_pointer = &codec;
      fd::DerivedTypeHolder<fec::Convolutional, fec::Turbo, fec::Ldpc> derived;
      // the archive fd::save writes (src/Serialization.h:92-100), into a stream of exact length
      // (fd::archiveSize only bounds it)
      std::ostringstream os;
      {
        boost::archive::binary_oarchive oa(os);
        derived.register_type(oa);
        oa & BOOST_SERIALIZATION_NVP(base_pointer);
      }
      std::string bytes = os.str();
      std::vector<char> check(fd::archiveSize(base_pointer, derived), 0);
      fd::save(base_pointer, check.data(), check.size(), derived);
      if (bytes.size() > check.size() || std::memcmp(bytes.data(), check.data(), bytes.size()) != 0)
        throw std::runtime_error("archive differs from fec::detail::save");
      FILE* f = fopen(argv[7], "wb");
      fwrite(bytes.data(), 1, bytes.size(), f);
      fclose(f);
      return 0;
    }
    if (mode == "load" && argc == 4) {
      FILE* f = fopen(argv[2], "rb");
      if (!f) throw std::runtime_error("cannot open archive");
      std::vector<char> buf;
      char tmp[65536];
      size_t n;
      while ((n = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
      fclose(f);
      fd::DerivedTypeHolder<fec::Convolutional, fec::Turbo, fec::Ldpc> derived;
      auto codec = fd::load<fec::Codec>(buf.data(), buf.size(), derived);
      auto* ldpc = dynamic_cast<fec::Ldpc*>(codec.get());
      if (!ldpc) throw std::runtime_error("archive does not hold an Ldpc codec");
      struct Peek : fec::Ldpc {
        static const fd::Ldpc::Structure& of(const fec::Ldpc& c) { return static_cast<const Peek&>(c).structure(); }
      };
      const auto& st = Peek::of(*ldpc);
      auto opt = ldpc->getDecoderOptions();
      FILE* o = fopen(argv[3], "wb");
      int64_t hdr[7] = {(int64_t)codec->msgSize(), (int64_t)codec->systSize(), (int64_t)codec->paritySize(),
                        (int64_t)codec->stateSize(), (int64_t)codec->getWorkGroupSize(), (int64_t)opt.algorithm(),
                        (int64_t)st.iterations()};
      fwrite(hdr, 8, 7, o);
      fclose(o);
      write_csr_rows(st.checks(), st.msgSize(), (std::string(argv[3]) + ".csr").c_str());
      return 0;
    }
    if (mode == "puncture" && argc == 6) {
      auto H = make_code(argv[2]);
      fd::Ldpc::Structure s{fec::Ldpc::EncoderOptions(H), make_options(2, 1, "1")};
      auto bits = [](const std::string& a) {
        std::vector<bool> v;
        if (a != "-")
          for (auto& t : split(a, ',')) v.push_back(t == "1");
        return v;
      };
      auto perm = s.puncturing(fec::Ldpc::PunctureOptions().mask(bits(argv[3])).systMask(bits(argv[4])));
      FILE* f = fopen(argv[5], "wb");
      int64_t hdr[2] = {(int64_t)perm.inputSize(), (int64_t)perm.outputSize()};
      fwrite(hdr, 8, 2, f);
      for (size_t i = 0; i < perm.outputSize(); ++i) { int64_t x = (int64_t)perm[i]; fwrite(&x, 8, 1, f); }
      fclose(f);
      return 0;
    }
    fprintf(stderr, "bad arguments\n");
    return 2;
  } catch (std::exception& e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
