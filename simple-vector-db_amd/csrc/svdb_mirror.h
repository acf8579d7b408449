// svdb_mirror.h — the host mirror's positional delete (internal).
#pragma once

#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <thread>

#include "vector_database.h"

namespace svdb {

// Removes v[index] from v[0..size) by moving the later records down one,
// as vector_database.c:189-191 does with one memmove.  Large moves (a
// delete near the front of a store of millions of rows moves hundreds of
// MB) are split over up to max_threads threads: each moves its own span,
// and the one record that crosses into the previous span is saved before
// any span moves.  Allocates nothing that can throw; a thread that cannot
// be started leaves its span to the calling thread.
inline void shift_down_one(Vector* v, size_t index, size_t size,
                           size_t parallel_bytes = (size_t)32 << 20, unsigned max_threads = 8) {
  if (index + 1 > size) return;
  const size_t n = size - index - 1;  // records that move
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t T = std::min<size_t>({8, max_threads, hw, n});
  if (n * sizeof(Vector) < parallel_bytes || T < 2) {
    memmove(&v[index], &v[index + 1], n * sizeof(Vector));
    return;
  }
  // span j: destination records [b_j, e_j) from [b_j + 1, e_j + 1); record
  // e_j, read by span j, is the first one span j + 1 overwrites
  Vector* base = v + index;
  size_t b[8], e[8];
  Vector edge[8];
  for (size_t j = 0; j < T; j++) {
    b[j] = n * j / T;
    e[j] = n * (j + 1) / T;
    edge[j] = base[e[j]];  // e[j] <= n: inside v[index .. size)
  }
  auto span = [&](size_t j) {
    if (e[j] > b[j]) {
      memmove(&base[b[j]], &base[b[j] + 1], (e[j] - b[j] - 1) * sizeof(Vector));
      base[e[j] - 1] = edge[j];
    }
  };
  std::thread th[8];
  size_t started = 0;
  try {
    for (size_t j = 1; j < T; j++) {
      th[j] = std::thread(span, j);
      started = j;
    }
  } catch (...) {
  }
  span(0);
  for (size_t j = started + 1; j < T; j++) span(j);
  for (size_t j = 1; j <= started; j++) th[j].join();
}

}  // namespace svdb
