// Device-wide inclusive prefix sum of uint64 counts (hipCUB), used to turn
// per-query hit counts into CSR offsets.
#include <hipcub/hipcub.hpp>

#include "kernels.hpp"

namespace hibf_amd
{

hipError_t inclusive_sum_u64(void * temp, size_t & temp_bytes, uint64_t const * in, uint64_t * out, uint64_t n,
                             hipStream_t stream)
{
    return hipcub::DeviceScan::InclusiveSum(temp, temp_bytes, in, out, static_cast<int>(n), stream);
}

} // namespace hibf_amd
