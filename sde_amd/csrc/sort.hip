// Key/index radix sort of the visibilities (rocPRIM's device radix sort).
//
// The cell sort orders the visibilities by tile-major cell key; the values
// are the original index and the weight of each point.  The sort is
// stable, so points of one cell keep their original order and the sorted
// order -- and every result built from it -- is deterministic.  Only bits
// [0, nbits) order the keys; bits above ride along (bit 31 carries a flag
// of the prep pass).
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "internal.h"

size_t sort_temp_bytes(int64_t n, int nbits) {
    size_t bytes = 0;
    if (rocprim::radix_sort_pairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                  (const uint64_t *)nullptr, (uint64_t *)nullptr, (size_t)n, 0u,
                                  (unsigned)nbits) != hipSuccess)
        return 0;
    return bytes;
}

hipError_t launch_radix_sort(hipStream_t s, void *temp, size_t temp_bytes, const uint32_t *keys_in,
                             uint32_t *keys_out, const uint64_t *vals_in, uint64_t *vals_out, int64_t n,
                             int nbits) {
    if (n <= 0) return hipSuccess;
    size_t bytes = temp_bytes;
    return rocprim::radix_sort_pairs(temp, bytes, keys_in, keys_out, vals_in, vals_out, (size_t)n, 0u,
                                     (unsigned)nbits, s);
}
