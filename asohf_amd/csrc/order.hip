// order.hip -- particle orderings for gather kernels: a radix sort of
// (base-cell key, particle index) pairs with rocPRIM, kept in its own
// translation unit (the library headers are large).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "launch.h"

namespace asohf {

hipError_t sort_cell_pairs(void* temp, size_t& temp_bytes, const unsigned* key_in, unsigned* key_out,
                           const unsigned* idx_in, unsigned* idx_out, unsigned long long n, int end_bit,
                           hipStream_t s) {
  return rocprim::radix_sort_pairs(temp, temp_bytes, key_in, key_out, idx_in, idx_out, n, 0u,
                                   (unsigned)end_bit, s);
}

}  // namespace asohf
