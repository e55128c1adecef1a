// Device scan and radix sort used by the BVH builder.
#pragma once
#include <vector>

#include "gpu_util.h"

namespace rtg {

// Exclusive prefix sum of uint32 values (multi-level tile scan).
struct Scanner {
  DevBuf tmp;
  void exclusive(const uint32_t* in, uint32_t* out, size_t n, hipStream_t s);
};

// Stable LSD radix sort of (key, value) pairs over bits [begin_bit, end_bit)
// in 8-bit digits; the result is left in keys/vals.
struct RadixSorter {
  DevBuf hist, offs, kt, vt;
  Scanner scanner;
  void sort_pairs(uint32_t* keys, uint32_t* vals, size_t n, int begin_bit, int end_bit, hipStream_t s);
};

}  // namespace rtg
