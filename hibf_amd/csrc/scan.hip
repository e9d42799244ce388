// Device-wide hipCUB primitives: the inclusive prefix sum that turns
// per-query hit counts into CSR offsets, and the pair sort of the tracked
// emplace.
#include <hipcub/hipcub.hpp>

#include "kernels.hpp"

namespace hibf_amd
{

hipError_t inclusive_sum_u64(void * temp, size_t & temp_bytes, uint64_t const * in, uint64_t * out, uint64_t n,
                             hipStream_t stream)
{
    return hipcub::DeviceScan::InclusiveSum(temp, temp_bytes, in, out, static_cast<int>(n), stream);
}

// hipCUB stable radix sort of (key, value) pairs on key bits [0, end_bit);
// call with temp == nullptr to size temp_bytes.
hipError_t radix_sort_pairs_u64(void * temp, size_t & temp_bytes, uint64_t const * keys_in, uint64_t * keys_out,
                                uint64_t const * vals_in, uint64_t * vals_out, uint64_t n, int end_bit,
                                hipStream_t stream)
{
    return hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_in, keys_out, vals_in, vals_out,
                                              static_cast<int>(n), 0, end_bit, stream);
}

} // namespace hibf_amd
