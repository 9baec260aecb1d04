// Device-side bit helpers on one-word occupation strings, shared by the HIP sources.
#pragma once

#include "pqg_ci.hpp"

namespace pqg {

__device__ __forceinline__ int ctz64(uint64_t x) { return __builtin_ctzll(x); }
__device__ __forceinline__ uint64_t low_mask(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }
// bits strictly between lo and hi
__device__ __forceinline__ uint64_t between_mask(int lo, int hi) { return low_mask(hi) & ~low_mask(lo + 1); }

// Index of string J in a set (internal order): colex rank through the per-byte rank table,
// then the rank -> index map of a truncated set (-1 when J is not in the set).
__device__ __forceinline__ int str_index(const SlotDev& s, const uint32_t* tab, uint64_t J) {
  if (s.nocc == 1 && !s.r2i) return __builtin_ctzll(J);  // colex rank of {p} is p
  uint32_t r = 0;
  int e = s.nocc;
  const int w = s.nocc + 1;
  for (int c = s.nchunk - 1; c >= 0; --c) {
    const uint32_t x = (uint32_t)(J >> (8 * c)) & 255u;
    r += tab[((c << 8) + x) * w + e];
    e -= __popc(x);
  }
  return s.r2i ? s.r2i[r] : (int)r;
}

// Position of the k-th set bit of x (k counted from 0, ascending).
__device__ __forceinline__ int nth_bit(uint64_t x, int k) {
  for (int i = 0; i < k; i++) x &= x - 1;
  return ctz64(x);
}

}  // namespace pqg
