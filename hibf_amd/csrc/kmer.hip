// k-mer hashes of DNA reads on the device: the values the reference's tutorial
// search hands to membership_for (test/tutorial/search.cpp:37-41), there
// computed on the host by seqan3::views::kmer_hash(seqan3::ungapped{k}) over
// dna4 sequences.  The hash of a window is the base-4 number of its letters'
// dna4 ranks, first letter most significant (A=0, C=1, G=2, T=U=3; dna4 reads
// every other character as A).  A read of length n < k has no k-mers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>

#include "hibf_gpu.h"
#include "kernels.hpp"

namespace hibf_amd
{
int set_error(int status, std::string msg);
}

namespace
{

__device__ __forceinline__ uint64_t dna4_rank(uint8_t c)
{
    switch (c | 0x20u) // lower case
    {
    case 'c': return 1u;
    case 'g': return 2u;
    case 't':
    case 'u': return 3u;
    default: return 0u; // 'a' and everything dna4 maps to A
    }
}

__global__ void __launch_bounds__(256) kmer_count_kernel(uint64_t const * __restrict__ seq_offsets,
                                                         uint64_t n,
                                                         uint32_t k,
                                                         uint64_t * __restrict__ counts)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    {
        uint64_t const len = seq_offsets[i + 1] - seq_offsets[i];
        counts[i] = len >= k ? len - k + 1u : 0u;
    }
}

__global__ void zero_first_kernel(uint64_t * p)
{
    if (threadIdx.x == 0)
        p[0] = 0u;
}

// One block of 256 threads per read (grid-stride).  The read is taken in
// segments of up to 8,192 bases that overlap by k - 1: the block stages a
// segment's bytes in LDS (coalesced), packs them into 2-bit ranks, 32 bases per
// 64-bit word with the first base most significant, and every window's hash
// is then a funnel shift of two neighbouring words: bits 2 p .. 2 p + 2 k of the
// packed stream.
constexpr uint32_t kHashThreads = 256;
constexpr uint32_t kSegment = 32u * kHashThreads;

__global__ void __launch_bounds__(kHashThreads) kmer_hash_kernel(uint8_t const * __restrict__ seq,
                                                                 uint64_t const * __restrict__ seq_offsets,
                                                                 uint64_t n,
                                                                 uint32_t k,
                                                                 uint64_t const * __restrict__ hash_offsets,
                                                                 uint64_t * __restrict__ hashes)
{
    __shared__ uint8_t raw[kSegment];
    __shared__ uint64_t words[kHashThreads + 1];
    uint32_t const tid = threadIdx.x;
    for (uint64_t i = blockIdx.x; i < n; i += gridDim.x)
    {
        uint8_t const * const s = seq + seq_offsets[i];
        uint64_t const len = seq_offsets[i + 1] - seq_offsets[i];
        uint64_t const first = hash_offsets[i];
        uint64_t const count = hash_offsets[i + 1] - first;
        for (uint64_t b0 = 0; b0 < count; b0 += kSegment - (k - 1u))
        {
            uint32_t const seg = static_cast<uint32_t>(min<uint64_t>(kSegment, len - b0));
            uint32_t const nwords = (seg + 31u) / 32u;
            for (uint32_t x = tid; x < nwords * 32u; x += kHashThreads)
                raw[x] = x < seg ? s[b0 + x] : static_cast<uint8_t>('A');
            __syncthreads();
            if (tid < nwords)
            {
                uint64_t w = 0;
#pragma unroll 8
                for (uint32_t j = 0; j < 32u; ++j)
                    w = (w << 2) | dna4_rank(raw[32u * tid + j]);
                words[tid] = w;
            }
            if (tid == 0)
                words[nwords] = 0;
            __syncthreads();
            uint64_t const end = min<uint64_t>(count, b0 + seg - (k - 1u));
            for (uint64_t p = b0 + tid; p < end; p += kHashThreads)
            {
                uint32_t const q = static_cast<uint32_t>(p - b0);
                uint32_t const off = q & 31u;
                uint64_t const hi = words[q >> 5];
                uint64_t const stream = off ? (hi << (2u * off)) | (words[(q >> 5) + 1u] >> (64u - 2u * off)) : hi;
                hashes[first + p] = stream >> (64u - 2u * k);
            }
            __syncthreads();
        }
    }
}

// threshold = uint16_t(fraction * number of k-mers) (test/tutorial/search.cpp:42),
// truncated toward zero and capped at 65535.
__global__ void __launch_bounds__(256) kmer_threshold_kernel(uint64_t const * __restrict__ hash_offsets,
                                                             uint64_t n,
                                                             double fraction,
                                                             uint16_t * __restrict__ thresholds)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    {
        double const t = fraction * static_cast<double>(hash_offsets[i + 1] - hash_offsets[i]);
        thresholds[i] = t >= 65535.0 ? uint16_t{65535} : t <= 0.0 ? uint16_t{0} : static_cast<uint16_t>(t);
    }
}

unsigned blocks_for(uint64_t n, unsigned per)
{
    uint64_t const b = (n + per - 1u) / per;
    return static_cast<unsigned>(std::max<uint64_t>(1u, std::min<uint64_t>(b, 1u << 20)));
}

int device_error(hipError_t e, char const * what)
{
    (void)hipGetLastError();
    return hibf_amd::set_error(HIBF_GPU_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

} // namespace

namespace hibf_amd
{

// The hashes alone, with hash offsets the caller computed (n + 1 values,
// hash_offsets[i + 1] - hash_offsets[i] = max(0, length of read i - k + 1)).
hipError_t kmer_hash_given_offsets_async(uint8_t const * d_seq,
                                         uint64_t const * d_seq_offsets,
                                         uint64_t n,
                                         uint32_t k,
                                         uint64_t const * d_hash_offsets,
                                         uint64_t * d_hashes,
                                         hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(kmer_hash_kernel, dim3(blocks_for(n, 1)), dim3(kHashThreads), 0, s, d_seq, d_seq_offsets, n,
                       k, d_hash_offsets, d_hashes);
    return hipGetLastError();
}

} // namespace hibf_amd

extern "C" {

int hibf_gpu_kmer_hash(uint8_t const * d_seq,
                       uint64_t const * d_seq_offsets,
                       uint64_t n_seqs,
                       uint32_t k,
                       uint64_t * d_hash_offsets,
                       uint64_t * d_hashes,
                       uint64_t hashes_cap,
                       uint64_t * h_total,
                       void * stream)
{
    if (!d_seq_offsets || !d_hash_offsets || !h_total || (!d_seq && n_seqs) || (!d_hashes && hashes_cap))
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (k == 0 || k > 32)
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "k must be in [1, 32] for 64-bit k-mer hashes");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    *h_total = 0;
    hipLaunchKernelGGL(zero_first_kernel, dim3(1), dim3(64), 0, s, d_hash_offsets);
    if (n_seqs == 0)
    {
        hipError_t const e = hipStreamSynchronize(s);
        return e == hipSuccess ? HIBF_GPU_OK : device_error(e, "kmer_hash");
    }
    hipLaunchKernelGGL(kmer_count_kernel, dim3(blocks_for(n_seqs, 256)), dim3(256), 0, s, d_seq_offsets, n_seqs, k,
                       d_hash_offsets + 1);
    size_t temp_bytes = 0;
    hipError_t e = hibf_amd::inclusive_sum_u64(nullptr, temp_bytes, d_hash_offsets + 1, d_hash_offsets + 1, n_seqs, s);
    void * temp = nullptr;
    if (e == hipSuccess)
        e = hibf_amd::device_malloc(&temp, std::max<size_t>(temp_bytes, 1));
    if (e == hipSuccess)
        e = hibf_amd::inclusive_sum_u64(temp, temp_bytes, d_hash_offsets + 1, d_hash_offsets + 1, n_seqs, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_total, d_hash_offsets + n_seqs, 8u, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (temp)
        (void)hipFree(temp);
    if (e != hipSuccess)
        return device_error(e, "kmer_hash offsets");
    if (*h_total > hashes_cap)
        return hibf_amd::set_error(HIBF_GPU_ERR_CAPACITY, "hash buffer too small (see *h_total)");
    hipLaunchKernelGGL(kmer_hash_kernel, dim3(blocks_for(n_seqs, 1)), dim3(kHashThreads), 0, s, d_seq, d_seq_offsets,
                       n_seqs, k, d_hash_offsets, d_hashes);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    return e == hipSuccess ? HIBF_GPU_OK : device_error(e, "kmer_hash");
}

int hibf_gpu_kmer_thresholds(uint64_t const * d_hash_offsets,
                             uint64_t n_seqs,
                             double fraction,
                             uint16_t * d_thresholds,
                             void * stream)
{
    if (!d_hash_offsets || (!d_thresholds && n_seqs))
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (!(fraction >= 0.0))
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "fraction must be >= 0");
    if (n_seqs == 0)
        return HIBF_GPU_OK;
    hipLaunchKernelGGL(kmer_threshold_kernel, dim3(blocks_for(n_seqs, 256)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), d_hash_offsets, n_seqs, fraction, d_thresholds);
    hipError_t const e = hipGetLastError();
    return e == hipSuccess ? HIBF_GPU_OK : device_error(e, "kmer_thresholds");
}

} // extern "C"
