// kern_sort.hip -- particle sort by cell with compaction
// (particle_base::sort_by_cell, cuda/core/particle_base_impl.cu:435-464,
// and its CPU twin core/particle_base_impl.hpp:250-305): a stable sort on
// the key cell (empty slots keyed past the last cell), a gather of every
// SoA field, and number := count of non-empty slots.  The radix pass is
// rocPRIM's stable LSD sort through hipCUB; the key width is trimmed to
// the cell count, so a 2048^2 grid takes 23 bits.  Stability makes the
// order within a cell the arrival order, exactly like the reference's
// counting sort.
#include <hipcub/hipcub.hpp>

#include "apt_common.hpp"
#include "apt_sort.hpp"

namespace apt {

__global__ void k_make_keys(const uint32_t *__restrict__ cell, uint64_t n,
                            uint32_t empty_key, const uint32_t *tile_key_lut,
                            uint32_t *__restrict__ keys,
                            uint32_t *__restrict__ perm) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t c = cell[i];
  uint32_t k;
  if (c == MAX_CELL)
    k = empty_key;
  else
    k = tile_key_lut ? tile_key_lut[c] : c;
  keys[i] = k;
  perm[i] = (uint32_t)i;
}

// number of keys below empty_key in a sorted key array
__global__ void k_count_valid(const uint32_t *__restrict__ keys, uint64_t n,
                              uint32_t empty_key,
                              unsigned long long *__restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool here = keys[i] < empty_key;
  bool next = (i + 1 < n) ? keys[i + 1] < empty_key : false;
  if (i == 0 && !here) *out = 0;
  if (here && !next) *out = i + 1;
}

template <typename T>
__global__ void k_gather(const T *__restrict__ src,
                         const uint32_t *__restrict__ perm,
                         T *__restrict__ dst, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  dst[i] = src[perm[i]];
}

static inline unsigned blocks_for(uint64_t n) {
  return (unsigned)((n + 255) / 256);
}

int key_bits(uint32_t max_key) {
  int b = 1;
  while (b < 32 && (1ull << b) <= (uint64_t)max_key) b++;
  return b;
}

hipError_t SortWork::ensure(uint64_t n) {
  if (n <= cap) return hipSuccess;
  release();
  hipError_t e;
  if ((e = hipMalloc(&keys_in, n * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(&keys_out, n * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(&perm_in, n * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(&perm_out, n * 4)) != hipSuccess) return e;
  if ((e = hipMalloc(&count, sizeof(unsigned long long))) != hipSuccess)
    return e;
  size_t bytes = 0;
  e = hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out,
                                         perm_in, perm_out, (int)n, 0, 32);
  if (e != hipSuccess) return e;
  if ((e = hipMalloc(&temp, bytes)) != hipSuccess) return e;
  temp_bytes = bytes;
  cap = n;
  return hipSuccess;
}

void SortWork::release() {
  (void)hipFree(keys_in);
  (void)hipFree(keys_out);
  (void)hipFree(perm_in);
  (void)hipFree(perm_out);
  (void)hipFree(count);
  (void)hipFree(temp);
  keys_in = keys_out = perm_in = perm_out = nullptr;
  count = nullptr;
  temp = nullptr;
  temp_bytes = 0;
  cap = 0;
}

hipError_t sort_keys(SortWork &w, const uint32_t *cell, uint64_t n,
                     uint32_t empty_key, const uint32_t *tile_key_lut,
                     hipStream_t st) {
  hipError_t e = w.ensure(n);
  if (e != hipSuccess) return e;
  k_make_keys<<<blocks_for(n), 256, 0, st>>>(cell, n, empty_key, tile_key_lut,
                                             w.keys_in, w.perm_in);
  size_t bytes = w.temp_bytes;
  e = hipcub::DeviceRadixSort::SortPairs(w.temp, bytes, w.keys_in,
                                         w.keys_out, w.perm_in, w.perm_out,
                                         (int)n, 0, key_bits(empty_key), st);
  if (e != hipSuccess) return e;
  k_count_valid<<<blocks_for(n), 256, 0, st>>>(w.keys_out, n, empty_key,
                                               w.count);
  return hipGetLastError();
}

hipError_t hipcub_sort_pairs(SortWork &w, uint64_t n, int bits,
                             hipStream_t st) {
  size_t bytes = w.temp_bytes;
  return hipcub::DeviceRadixSort::SortPairs(w.temp, bytes, w.keys_in,
                                            w.keys_out, w.perm_in,
                                            w.perm_out, (int)n, 0, bits, st);
}

template <typename T>
hipError_t gather_field(SortWork &w, const T *src, T *dst, uint64_t n,
                        hipStream_t st) {
  k_gather<T><<<blocks_for(n), 256, 0, st>>>(src, w.perm_out, dst, n);
  return hipGetLastError();
}

template hipError_t gather_field<uint32_t>(SortWork &, const uint32_t *,
                                           uint32_t *, uint64_t, hipStream_t);
template hipError_t gather_field<float>(SortWork &, const float *, float *,
                                        uint64_t, hipStream_t);
template hipError_t gather_field<double>(SortWork &, const double *,
                                         double *, uint64_t, hipStream_t);
template hipError_t gather_field<uint64_t>(SortWork &, const uint64_t *,
                                           uint64_t *, uint64_t,
                                           hipStream_t);

}  // namespace apt
