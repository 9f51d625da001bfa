// Radix sort of (vertex, set) pairs for the selector's inverted index
// (select.hip): the pairs of a sparse pool sorted by vertex are the index
// itself, with no atomic per element.  rocPRIM's device radix sort, over
// the key bits the graph's vertex ids use.
#include <rocprim/device/device_radix_sort.hpp>

#include "internal.h"

namespace rip {

int sort_pairs_u32(uint32_t *k0, uint32_t *k1, uint32_t *v0, uint32_t *v1, uint64_t count,
                   uint32_t bits, void **temp, size_t *temp_bytes, hipStream_t st) {
  rocprim::double_buffer<uint32_t> keys(k0, k1), vals(v0, v1);
  size_t need = 0;
  RIP_HIP(rocprim::radix_sort_pairs(nullptr, need, keys, vals, static_cast<size_t>(count), 0u,
                                    bits, st));
  if (*temp_bytes < need) {
    dev_free(*temp);
    *temp = nullptr;
    *temp_bytes = 0;
    *temp = dev_alloc<uint8_t>(need);
    *temp_bytes = need;
  }
  size_t have = *temp_bytes;
  RIP_HIP(rocprim::radix_sort_pairs(*temp, have, keys, vals, static_cast<size_t>(count), 0u,
                                    bits, st));
  return keys.current() == k0 ? 0 : 1;
}

}  // namespace rip
