// FeCl's saved-codec format for fec::Ldpc (host side).
//
// FeCl saves a codec with Boost.Serialization 1.58's binary_oarchive, through a `const
// fec::Codec*` with the types Convolutional, Turbo and Ldpc registered in that order
// (src/Serialization.h:81-101; the Matlab wrapper's registration, wrap/wrap.h:35).  For an Ldpc
// codec that archive is a fixed sequence of fields (src/Codec.h:309-314, src/Ldpc.h:119-124,
// src/detail/Codec.h:186-193, src/detail/Ldpc.h:148-159, src/BitMatrix.h:53-57, 614-620): the
// work group size, the sizes and algorithm, the five sparse matrices of the structure
// (H_, DC_, T_, A_, B_), the iteration count and the scaling table.  This file writes and reads
// exactly that sequence; Boost itself is not used.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "ldpc_code.h"

namespace fecl {

struct ArchiveOptions {
  int32_t algorithm = 2;          // DecoderAlgorithm: 0 Exact, 1 Linear, 2 Approximate
  int64_t iterations = 0;
  std::vector<double> scaling;    // [rows][cols]
  int64_t rows = 0, cols = 0;
  int32_t work_group_size = 8;
};

// Appends the archive of the codec (structure g with E edges, options o) to *out.
void archive_write(const GeneratorForm& g, int64_t E, const ArchiveOptions& o, std::vector<uint8_t>* out);

// Parses an archive.  Bytes after the archive are ignored (FeCl's Matlab wrapper saves into a
// buffer with slack).  Returns "" or why the bytes are not an Ldpc codec archive.
std::string archive_read(const uint8_t* buf, int64_t size, GeneratorForm* g, ArchiveOptions* o);

}  // namespace fecl
