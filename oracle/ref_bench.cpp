// Times stock FeCl's public decode path (fec::Ldpc::decode, std::thread fan-out with
// workGroupSize = threads) on an LLR file.  Used as the CPU baseline beside the HIP decoder.
//   fecl_bench <code> <llr> <alg> <iters> <threads>
//   code = dvbs2:<n>:<rate> | gallager:<n>:<wc>:<wr>:<seed>
// Prints one JSON line: {"seconds": s, "frames": F, "threads": T, "k": K, "errors_vs_zero": e}
#include <chrono>
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>
#include <stdexcept>

#include "Ldpc.h"

static std::vector<std::string> split(const std::string& s, char c) {
  std::vector<std::string> out; std::string cur;
  for (char ch : s) { if (ch == c) { out.push_back(cur); cur.clear(); } else cur += ch; }
  out.push_back(cur); return out;
}

int main(int argc, char** argv) {
  try {
    if (argc != 6) { fprintf(stderr, "usage: fecl_bench <code> <llr> <alg> <iters> <threads>\n"); return 2; }
    auto p = split(argv[1], ':');
    fec::SparseBitMatrix H = p[0] == "dvbs2"
        ? fec::Ldpc::DvbS2::matrix(std::stoul(p[1]), std::stod(p[2]))
        : fec::Ldpc::Gallager::matrix(std::stoul(p[1]), std::stoul(p[2]), std::stoul(p[3]), std::stoull(p[4]));
    FILE* f = fopen(argv[2], "rb");
    if (!f) throw std::runtime_error("cannot open llr file");
    int64_t hdr[2];
    if (fread(hdr, 8, 2, f) != 2) throw std::runtime_error("short llr file");
    std::vector<double> llr(hdr[0] * hdr[1]);
    if (fread(llr.data(), 8, llr.size(), f) != llr.size()) throw std::runtime_error("short llr file");
    fclose(f);
    int threads = atoi(argv[5]);
    auto opt = fec::Ldpc::DecoderOptions().algorithm(fec::DecoderAlgorithm(atoi(argv[3]))).iterations(std::stoul(argv[4]));
    fec::Ldpc codec(fec::Ldpc::EncoderOptions(H), opt, threads);
    std::vector<fec::BitField<size_t>> msg;
    auto t0 = std::chrono::steady_clock::now();
    codec.decode(llr, msg);
    auto t1 = std::chrono::steady_clock::now();
    size_t ones = 0;
    for (auto& b : msg) ones += b ? 1 : 0;
    printf("{\"seconds\": %.6f, \"frames\": %lld, \"threads\": %d, \"k\": %zu, \"ones\": %zu}\n",
           std::chrono::duration<double>(t1 - t0).count(), (long long)hdr[0], threads, codec.msgSize(), ones);
    return 0;
  } catch (std::exception& e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
