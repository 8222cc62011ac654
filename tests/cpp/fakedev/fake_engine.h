// CPU fake of the decode engine behind the serving loop (see fake_engine.cpp).  A frame's
// "decode" is fixed by its channel values: hard bit v = (llr[v] < 0), and it finishes after
// fake_iterations(llr, I) steps, so a test knows every result without a decoder.
#pragma once
#include <stdint.h>

#include <cstring>

inline int fake_iterations(const float* llr, int iterations) {
  uint32_t a, b;
  std::memcpy(&a, &llr[0], 4);
  std::memcpy(&b, &llr[1], 4);
  const int t = 1 + (int)(((a * 2654435761u) ^ (b >> 7)) % 9u);
  return t < iterations ? t : iterations;
}
