// Which tiles of a serving pool take a round's admissions, in order (the serving loop in
// ldpc_server.cpp fills them lowest slot first until the round's frames are placed).  Pure
// host code, so the unit tests check it directly.
//
// A tile with a fresh frame reads the channel values on its next check update besides its
// messages, and an admission writes each frame's values across all of its tile's rows, so a
// round's admissions go to few tiles: at most `max_occupied` tiles that hold frames, then the
// tiles that hold none in index order.  First comes the fullest tile holding frames that has
// room for all `want` frames of the round, then the tiles holding frames with the most free
// slots.  Below saturation a round admits a few frames and they fill up nearly full tiles, so
// sparse tiles drain and stop costing steps; a saturated pool's round admits more than any
// tile has free, and its frames go to the tiles with the most room.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace fecl {

// free_slots[t] / used_slots[t]: free and occupied slots of tile t; order receives the tiles.
inline void admission_order(const std::vector<int64_t>& free_slots, const std::vector<int64_t>& used_slots,
                            int64_t want, int max_occupied, std::vector<int>* order) {
  order->clear();
  const int T = (int)free_slots.size();
  for (int t = 0; t < T; ++t)
    if (used_slots[t] > 0 && free_slots[t] > 0) order->push_back(t);
  std::stable_sort(order->begin(), order->end(), [&](int x, int y) { return free_slots[x] > free_slots[y]; });
  int best = -1;   // the last (so the fullest) of the tiles with room for the whole round
  for (int i = 0; i < (int)order->size(); ++i)
    if (free_slots[(*order)[i]] >= want) best = i;
  if (best > 0) std::rotate(order->begin(), order->begin() + best, order->begin() + best + 1);
  if ((int)order->size() > max_occupied) order->resize(max_occupied);
  for (int t = 0; t < T; ++t)
    if (used_slots[t] == 0) order->push_back(t);
}

}  // namespace fecl
