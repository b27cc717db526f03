// apt_sort.hpp -- scratch and entry points of the particle sort.
#pragma once
#include "apt_common.hpp"

namespace apt {

struct SortWork {
  uint32_t *keys_in = nullptr, *keys_out = nullptr;
  uint32_t *perm_in = nullptr, *perm_out = nullptr;
  unsigned long long *count = nullptr;  // non-empty slots after the sort
  void *temp = nullptr;
  size_t temp_bytes = 0;
  uint64_t cap = 0;
  hipError_t ensure(uint64_t n);
  void release();
};

int key_bits(uint32_t max_key);

// Sort slot indices by key (cell, or tile_key_lut[cell] when given);
// empty slots take empty_key and sort last.  Leaves the permutation in
// w.perm_out and the non-empty count in *w.count (device).
hipError_t sort_keys(SortWork &w, const uint32_t *cell, uint64_t n,
                     uint32_t empty_key, const uint32_t *tile_key_lut,
                     hipStream_t st);

// radix sort of (keys_in, perm_in) -> (keys_out, perm_out) on `bits` bits
hipError_t hipcub_sort_pairs(SortWork &w, uint64_t n, int bits,
                             hipStream_t st);

// dst[i] = src[perm_out[i]]
template <typename T>
hipError_t gather_field(SortWork &w, const T *src, T *dst, uint64_t n,
                        hipStream_t st);

}  // namespace apt
