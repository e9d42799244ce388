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

#include "device_common.hpp"
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

// ---------------------------------------------------------------------------
// Minimiser hashes: seqan3::views::minimiser_hash(shape, window_size, seed)
// over dna4 (seqan3 3.4.2, the version the reference pins for its tutorial,
// cmake/package-lock.cmake:50-56).  Per k-mer position i (n = length - size + 1
// of them) the value is min(forward[i] ^ seed, reverse[i] ^ seed): forward is
// the base-4 number of the ranks at the shape's used positions, reverse the
// same shape laid over the window's reverse complement.  Of a window of
// W = window_size - size + 1 values (W = n for shorter reads) the first emits
// its rightmost minimum; then, as each value enters, the tracked minimum is
// re-taken (rightmost on ties) and emitted if its position just left, else
// the entering value is emitted and tracked if strictly smaller.
//
// The same rule as a chain over positions: from the first window's rightmost
// minimum p, the next emitted position is the first t in (p, p + W] with a
// strictly smaller value (it enters before p leaves), else - when p + W is a
// position - the rightmost minimum of (p, p + W] (taken as p leaves).  Every
// lane computes the next position of its own positions from LDS; lane 0 walks
// the chain.  One wave per read, the read in segments of kMiniSeg chain
// positions whose values are held with a look-ahead of W.

constexpr uint32_t kMiniThreads = 64;
constexpr uint32_t kMiniSeg = 512;
constexpr uint32_t kMiniEnd = 0xFFFFFFFFu;

struct MiniParams
{
    uint64_t shape; // seqan3's binary literal: most significant set bit = first position
    uint64_t seed;
    uint32_t size;  // bit length of shape, <= 32
    uint32_t window; // W, values per window, <= kMaxMinimiserWindow
    uint32_t gapped;
    uint32_t words_cap; // packed 64-bit words of bases per segment (+ one zero word)
};

// Reverse complement of a window of `size` bases packed 2 bits each (first base
// most significant): the digits in reverse order, each r -> 3 - r.
__device__ __forceinline__ uint64_t revcomp_window(uint64_t x, uint32_t size)
{
    uint64_t r = __builtin_bitreverse64(x);
    r = ((r >> 1) & 0x5555555555555555ull) | ((r & 0x5555555555555555ull) << 1);
    r >>= 64u - 2u * size;
    uint64_t const ones = size == 32u ? ~0ull : (1ull << (2u * size)) - 1ull;
    return r ^ ones;
}

// The ranks at the shape's used positions, first position most significant.
__device__ __forceinline__ uint64_t shape_gather(uint64_t x, uint64_t shape, uint32_t size)
{
    uint64_t h = 0;
    for (uint32_t t = 0; t < size; ++t)
        if ((shape >> (size - 1u - t)) & 1u)
            h = (h << 2) | ((x >> (2u * (size - 1u - t))) & 3u);
    return h;
}

size_t minimiser_lds_bytes(MiniParams const & mp)
{
    return (static_cast<size_t>(mp.words_cap) + kMiniSeg + mp.window) * 8u + kMiniSeg * 4u;
}

// COUNT: counts[i] = number of minimisers of read i.  WRITE: read i's
// minimisers to hashes[hash_offsets[i] ...].
template <bool WRITE>
__global__ void __launch_bounds__(kMiniThreads) minimiser_kernel(uint8_t const * __restrict__ seq,
                                                                 uint64_t const * __restrict__ seq_offsets,
                                                                 uint64_t n,
                                                                 MiniParams mp,
                                                                 uint64_t * __restrict__ counts,
                                                                 uint64_t const * __restrict__ hash_offsets,
                                                                 uint64_t * __restrict__ hashes)
{
    extern __shared__ uint64_t mini_lds[];
    uint64_t * const words = mini_lds;                         // words_cap
    uint64_t * const vals = mini_lds + mp.words_cap;           // kMiniSeg + W
    uint32_t * const nxt = reinterpret_cast<uint32_t *>(vals + kMiniSeg + mp.window); // kMiniSeg
    uint32_t const lane = threadIdx.x;
    uint32_t const size = mp.size;
    for (uint64_t i = blockIdx.x; i < n; i += gridDim.x)
    {
        uint8_t const * const s = seq + seq_offsets[i];
        uint64_t const len = seq_offsets[i + 1] - seq_offsets[i];
        uint64_t const nk = len >= size ? len - size + 1u : 0u;
        uint64_t emitted = 0;
        uint64_t const out = WRITE ? hash_offsets[i] : 0u;
        uint64_t const we = min<uint64_t>(mp.window, nk);
        uint64_t p = 0;
        bool live = nk != 0;
        for (uint64_t b0 = 0; live && b0 < nk; b0 += kMiniSeg)
        {
            uint64_t const e = min<uint64_t>(nk, b0 + kMiniSeg + we); // values held: [b0, e)
            uint32_t const m = static_cast<uint32_t>(e - b0);
            uint32_t const nw = (m + size - 1u + 31u) / 32u;           // words of bases [b0, e + size - 1)
            for (uint32_t x = lane; x <= nw; x += kMiniThreads)
            {
                uint64_t w = 0;
                if (x < nw)
                {
                    uint64_t const first = b0 + 32u * x;
                    for (uint32_t j = 0; j < 32u; ++j)
                        w = (w << 2) | (first + j < len ? dna4_rank(s[first + j]) : 0u);
                }
                words[x] = w;
            }
            __syncthreads();
            for (uint32_t q = lane; q < m; q += kMiniThreads)
            {
                uint32_t const off = q & 31u;
                uint64_t const hi = words[q >> 5];
                uint64_t const stream = off ? (hi << (2u * off)) | (words[(q >> 5) + 1u] >> (64u - 2u * off)) : hi;
                uint64_t const win = stream >> (64u - 2u * size);
                uint64_t const rc = revcomp_window(win, size);
                uint64_t const fwd = mp.gapped ? shape_gather(win, mp.shape, size) : win;
                uint64_t const rev = mp.gapped ? shape_gather(rc, mp.shape, size) : rc;
                vals[q] = min(fwd ^ mp.seed, rev ^ mp.seed);
            }
            __syncthreads();
            uint32_t const pe = static_cast<uint32_t>(min<uint64_t>(nk, b0 + kMiniSeg) - b0);
            for (uint32_t q = lane; q < pe; q += kMiniThreads)
            {
                uint64_t const v = vals[q];
                uint64_t const pa = b0 + q;
                uint64_t const last = min<uint64_t>(pa + we, nk - 1u);
                uint32_t nx = kMiniEnd;
                uint64_t best = 0;
                uint32_t best_at = kMiniEnd;
                for (uint64_t t = pa + 1u; t <= last; ++t)
                {
                    uint64_t const x = vals[t - b0];
                    if (x < v)
                    {
                        nx = static_cast<uint32_t>(t - b0);
                        break;
                    }
                    if (best_at == kMiniEnd || x <= best)
                    {
                        best = x;
                        best_at = static_cast<uint32_t>(t - b0);
                    }
                }
                if (nx == kMiniEnd && pa + we <= nk - 1u)
                    nx = best_at;
                nxt[q] = nx;
            }
            if (b0 == 0)
            {
                // The first window's rightmost minimum: per lane, then across
                // the wave (smaller value, or the same value further right).
                uint64_t best = ~0ull;
                uint32_t best_at = 0;
                for (uint32_t q = lane; q < we; q += kMiniThreads)
                    if (vals[q] <= best)
                    {
                        best = vals[q];
                        best_at = q;
                    }
                for (int d = 32; d >= 1; d >>= 1)
                {
                    uint64_t const ob = __shfl_xor(static_cast<unsigned long long>(best), d);
                    uint32_t const oa = __shfl_xor(best_at, d);
                    if (ob < best || (ob == best && oa > best_at))
                    {
                        best = ob;
                        best_at = oa;
                    }
                }
                p = best_at;
                if (lane == 0)
                {
                    if (WRITE)
                        hashes[out] = best;
                    emitted = 1;
                }
            }
            __syncthreads();
            if (lane == 0)
            {
                while (p < b0 + pe)
                {
                    uint32_t const nx = nxt[p - b0];
                    if (nx == kMiniEnd)
                    {
                        live = false;
                        break;
                    }
                    if (WRITE)
                        hashes[out + emitted] = vals[nx];
                    ++emitted;
                    p = b0 + nx;
                }
            }
            p = hibf_amd::uniform64(p);
            live = hibf_amd::uniform32(live ? 1u : 0u) != 0u;
            __syncthreads(); // the next segment overwrites the LDS arrays
        }
        if (!WRITE && lane == 0)
            counts[i] = emitted;
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

unsigned minimiser_blocks(uint64_t n)
{
    return static_cast<unsigned>(std::max<uint64_t>(1u, std::min<uint64_t>(n, 1u << 16)));
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

int hibf_gpu_minimiser_hash(uint8_t const * d_seq,
                            uint64_t const * d_seq_offsets,
                            uint64_t n_seqs,
                            uint64_t shape,
                            uint32_t window_size,
                            uint64_t seed,
                            uint64_t * d_hash_offsets,
                            uint64_t * d_hashes,
                            uint64_t hashes_cap,
                            uint64_t * h_total,
                            void * stream)
{
    if (!d_seq_offsets || !d_hash_offsets || !h_total || (!d_seq && n_seqs) || (!d_hashes && hashes_cap))
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (shape == 0)
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "the shape must have at least one position");
    uint32_t const size = 64u - static_cast<uint32_t>(__builtin_clzll(shape));
    if (size > 32u)
        return hibf_amd::set_error(HIBF_GPU_ERR_UNSUPPORTED, "shapes of more than 32 positions are not supported");
    if (window_size < size)
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT,
                                   "The size of the shape cannot be greater than the window size.");
    uint64_t const w = static_cast<uint64_t>(window_size) - size + 1u;
    if (w > HIBF_GPU_MAX_MINIMISER_WINDOW)
        return hibf_amd::set_error(HIBF_GPU_ERR_UNSUPPORTED,
                                   "windows of more than 4096 k-mers (window_size - shape size + 1) are not supported");
    MiniParams mp{};
    mp.shape = shape;
    mp.seed = seed;
    mp.size = size;
    mp.window = static_cast<uint32_t>(w);
    mp.gapped = shape != ((size == 64u ? 0ull : (1ull << size)) - 1ull) ? 1u : 0u;
    mp.words_cap = (kMiniSeg + mp.window + size - 1u + 31u) / 32u + 1u;
    size_t const lds = minimiser_lds_bytes(mp);
    hipStream_t const s = static_cast<hipStream_t>(stream);
    *h_total = 0;
    hipLaunchKernelGGL(zero_first_kernel, dim3(1), dim3(64), 0, s, d_hash_offsets);
    if (n_seqs == 0)
    {
        hipError_t const e = hipStreamSynchronize(s);
        return e == hipSuccess ? HIBF_GPU_OK : device_error(e, "minimiser_hash");
    }
    hipLaunchKernelGGL(minimiser_kernel<false>, dim3(minimiser_blocks(n_seqs)), dim3(kMiniThreads), lds, s, d_seq,
                       d_seq_offsets, n_seqs, mp, d_hash_offsets + 1, nullptr, nullptr);
    hipError_t e = hipGetLastError();
    size_t temp_bytes = 0;
    if (e == hipSuccess)
        e = hibf_amd::inclusive_sum_u64(nullptr, temp_bytes, d_hash_offsets + 1, d_hash_offsets + 1, n_seqs, s);
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
        return device_error(e, "minimiser_hash offsets");
    if (*h_total > hashes_cap)
        return hibf_amd::set_error(HIBF_GPU_ERR_CAPACITY, "hash buffer too small (see *h_total)");
    hipLaunchKernelGGL(minimiser_kernel<true>, dim3(minimiser_blocks(n_seqs)), dim3(kMiniThreads), lds, s, d_seq,
                       d_seq_offsets, n_seqs, mp, nullptr, d_hash_offsets, d_hashes);
    e = hipGetLastError();
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    return e == hipSuccess ? HIBF_GPU_OK : device_error(e, "minimiser_hash");
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
