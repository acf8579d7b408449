// CPU check of svdb::shift_down_one (the host mirror's positional delete):
// for random sizes, positions and thread counts, the parallel split gives
// the array a plain memmove gives.  Built and run by test_mirror_cpu.py.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "svdb_mirror.h"

int main() {
  unsigned seed = 12345;
  auto rnd = [&]() { return seed = seed * 1103515245u + 12345u; };
  for (int trial = 0; trial < 400; trial++) {
    const size_t size = 1 + rnd() % 5000;
    const size_t index = rnd() % size;
    const unsigned threads = 1 + rnd() % 8;
    std::vector<Vector> a(size), ref(size);
    for (size_t i = 0; i < size; i++) {
      a[i].dimension = i;
      a[i].data = reinterpret_cast<double*>(i * 8 + 8);
      snprintf(a[i].uuid, UUID_SIZE, "u%zu", i);
    }
    ref = a;
    memmove(&ref[index], &ref[index + 1], (size - index - 1) * sizeof(Vector));
    svdb::shift_down_one(a.data(), index, size, /*parallel_bytes=*/0, threads);
    for (size_t i = 0; i + 1 < size; i++)
      if (a[i].dimension != ref[i].dimension || a[i].data != ref[i].data ||
          strcmp(a[i].uuid, ref[i].uuid) != 0) {
        fprintf(stderr, "trial %d size %zu index %zu threads %u: record %zu differs\n", trial,
                size, index, threads, i);
        return 1;
      }
  }
  puts("ok");
  return 0;
}
