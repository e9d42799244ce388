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
// position - the rightmost minimum of (p, p + W] (taken as p leaves).  One
// wave per read, the read in segments of kMiniSeg chain positions whose values
// are held with a look-ahead of W.

constexpr uint32_t kMiniThreads = 64; // one wave per read
constexpr uint32_t kMiniSeg = 512;

struct MiniParams
{
    uint64_t shape; // seqan3's binary literal: most significant set bit = first position
    uint64_t seed;
    uint32_t size;  // bit length of shape, <= 32
    uint32_t window; // W, values per window, <= HIBF_GPU_MAX_MINIMISER_WINDOW
    uint32_t gapped;
    uint32_t words_cap; // packed 64-bit words of bases per segment, whole 256-base chunks (+ one zero word)
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

// LDS of one block: the segment's packed bases (words_cap words), its values
// (kMiniSeg + W), next positions (kMiniSeg, 16-bit, from a 16-byte boundary)
// and chain flags (kMiniSeg bytes).
__host__ __device__ inline size_t minimiser_next_offset(MiniParams const & mp) // in 64-bit words
{
    return (static_cast<size_t>(mp.words_cap) + kMiniSeg + mp.window + 1u) & ~static_cast<size_t>(1);
}

size_t minimiser_lds_bytes(MiniParams const & mp)
{
    return minimiser_next_offset(mp) * 8u + kMiniSeg * 3u;
}

constexpr uint32_t kMiniSink = 0xFFFFFFFFu; // no next position: the chain ends
constexpr uint32_t kMiniOwn = kMiniSeg / kMiniThreads; // consecutive positions per lane (8)

// Next chain position after segment position q (absolute pa = b0 + q): the
// first of the following values (up to we of them, within the read) that is
// strictly smaller, else - when the window after pa is whole - the rightmost
// minimum of those values; kMiniSink when neither exists.
__device__ __forceinline__ uint32_t minimiser_next(uint64_t const * vals, uint32_t q, uint64_t pa, uint64_t we,
                                                   uint64_t nk)
{
    uint64_t const v = vals[q];
    uint32_t const cnt = static_cast<uint32_t>(min<uint64_t>(pa + we, nk - 1u) - pa); // values after pa to scan
    uint32_t nx = kMiniSink;
    uint32_t best_at = kMiniSink;
    uint64_t best = 0;
    // Four values per round trip: their loads are issued before any is compared.
    for (uint32_t t0 = 1; t0 <= cnt && nx == kMiniSink; t0 += 4u)
    {
        uint64_t x[4];
#pragma unroll
        for (uint32_t u = 0; u < 4u; ++u)
            x[u] = vals[q + (t0 + u <= cnt ? t0 + u : 0u)];
#pragma unroll
        for (uint32_t u = 0; u < 4u; ++u)
            if (t0 + u <= cnt && nx == kMiniSink)
            {
                if (x[u] < v)
                    nx = q + t0 + u;
                else if (best_at == kMiniSink || x[u] <= best)
                {
                    best = x[u];
                    best_at = q + t0 + u;
                }
            }
    }
    if (nx == kMiniSink && pa + we <= nk - 1u)
        nx = best_at;
    return nx;
}

// COUNT: counts[i] = number of minimisers of read i.  WRITE: read i's
// minimisers to hashes[hash_offsets[i] ...].
//
// Per segment: lanes pack the bases (4 per lane, 8 lanes joined by shuffles
// into a 64-bit word), hash every position into LDS and take every position's
// next chain position (neighbouring lanes on neighbouring positions).  Each
// lane then takes 8 consecutive positions and works out, for every entry into
// them, where the chain leaves them and which it visits (a bit mask); the
// wave follows the chain from lane to lane through those exits with readlane
// (scalar), each lane flags the positions visited from its entry, and the
// flagged values are written in position order (ballot prefix sums).
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
    uint64_t * const words = mini_lds;               // words_cap
    uint64_t * const vals = mini_lds + mp.words_cap; // kMiniSeg + W
    uint16_t * const next16 = reinterpret_cast<uint16_t *>(mini_lds + minimiser_next_offset(mp)); // kMiniSeg
    uint8_t * const flags = reinterpret_cast<uint8_t *>(next16 + kMiniSeg);               // kMiniSeg
    uint32_t const lane = threadIdx.x;
    uint32_t const size = mp.size;
    uint32_t const own = lane * kMiniOwn; // this lane's first position in a segment
    for (uint64_t i = blockIdx.x; i < n; i += gridDim.x)
    {
        uint8_t const * const s = seq + seq_offsets[i];
        uint64_t const len = seq_offsets[i + 1] - seq_offsets[i];
        uint64_t const nk = len >= size ? len - size + 1u : 0u;
        uint64_t emitted = 0;
        uint64_t const out = WRITE ? hash_offsets[i] : 0u;
        uint64_t const we = min<uint64_t>(mp.window, nk);
        uint64_t p = 0; // the chain's current position (absolute), emitted in the segment that holds it
        bool live = nk != 0;
        for (uint64_t b0 = 0; live && b0 < nk; b0 += kMiniSeg)
        {
            uint64_t const e = min<uint64_t>(nk, b0 + kMiniSeg + we); // values held: [b0, e)
            uint32_t const m = static_cast<uint32_t>(e - b0);
            uint32_t const chunks = (m + size - 1u + 255u) / 256u;      // 256 bases [b0 + 256 c, ...) each
            for (uint32_t c = 0; c < chunks; ++c)
            {
                uint64_t const at = b0 + 256u * c + 4u * lane;
                uint32_t v = 0;
#pragma unroll
                for (uint32_t j = 0; j < 4u; ++j)
                    v = (v << 2) | (at + j < len ? dna4_rank(s[at + j]) : 0u);
                unsigned long long x = v;
                x = (x << 8) | __shfl_down(x, 1);
                x = (x << 16) | __shfl_down(x, 2);
                x = (x << 32) | __shfl_down(x, 4);
                if ((lane & 7u) == 0u)
                    words[8u * c + (lane >> 3)] = x;
            }
            if (lane == 0)
                words[8u * chunks] = 0;
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
            }
            uint32_t const pe = static_cast<uint32_t>(min<uint64_t>(nk, b0 + kMiniSeg) - b0);
            if (p >= b0 + pe)
            {
                __syncthreads();
                continue; // the chain jumps past this segment's positions (large windows)
            }
            // Next positions (16-bit; beyond the segment's values: the sink),
            // then per lane 8 consecutive ones and per entry offset r the
            // positions visited (mask) and the exit (the first position after
            // them, in a later lane or segment, or the sink).
            for (uint32_t q = lane; q < pe; q += kMiniThreads)
            {
                uint32_t const nx = minimiser_next(vals, q, b0 + q, we, nk);
                next16[q] = static_cast<uint16_t>(nx == kMiniSink ? 0xFFFFu : nx);
            }
            __syncthreads();
            uint32_t nxt[kMiniOwn];
            {
                uint4 const packed = reinterpret_cast<uint4 const *>(next16)[lane]; // positions own .. own + 7
                uint32_t const w4[4] = {packed.x, packed.y, packed.z, packed.w};
#pragma unroll
                for (uint32_t r = 0; r < kMiniOwn; ++r)
                {
                    uint32_t const h = (w4[r / 2] >> (16u * (r % 2))) & 0xFFFFu;
                    nxt[r] = h == 0xFFFFu ? kMiniSink : h;
                }
            }
            uint32_t const stay = min(own + kMiniOwn, pe); // positions below this stay in the lane
            uint32_t mask[kMiniOwn];
            uint32_t exit[kMiniOwn];
#pragma unroll
            for (int r = static_cast<int>(kMiniOwn) - 1; r >= 0; --r)
            {
                uint32_t const nx = nxt[r];
                mask[r] = own + r < pe ? 1u << r : 0u;
                exit[r] = own + r < pe ? nx : kMiniSink;
                if (own + r < pe && nx != kMiniSink && nx < stay)
#pragma unroll
                    for (uint32_t j = r + 1; j < kMiniOwn; ++j)
                        if (nx == own + j)
                        {
                            mask[r] |= mask[j];
                            exit[r] = exit[j];
                        }
            }
            // Follow the chain across lanes (uniform): each lane's exits in
            // LDS (over its own next positions, which it alone read), then one
            // broadcast read per lane the chain passes through.
            {
                uint32_t h[kMiniOwn];
#pragma unroll
                for (uint32_t r = 0; r < kMiniOwn; ++r)
                    h[r] = exit[r] == kMiniSink ? 0xFFFFu : exit[r];
                reinterpret_cast<uint4 *>(next16)[lane] =
                    make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16));
            }
            __syncthreads();
            uint32_t entry = 0xFFu;
            uint32_t at = static_cast<uint32_t>(p - b0);
            uint32_t x = kMiniSink;
            for (;;)
            {
                entry = lane == at / kMiniOwn ? at % kMiniOwn : entry;
                uint32_t const h = hibf_amd::uniform32(next16[at]);
                x = h == 0xFFFFu ? kMiniSink : h;
                if (x == kMiniSink || x >= pe)
                    break;
                at = x;
            }
            uint32_t mine = 0;
#pragma unroll
            for (uint32_t r = 0; r < kMiniOwn; ++r)
                if (entry == r)
                    mine = mask[r];
            if (WRITE)
            {
                uint64_t bytes = 0;
#pragma unroll
                for (uint32_t r = 0; r < kMiniOwn; ++r)
                    bytes |= static_cast<uint64_t>((mine >> r) & 1u) << (8u * r);
                reinterpret_cast<uint64_t *>(flags)[lane] = bytes;
                __syncthreads();
                for (uint32_t c0 = 0; c0 < pe; c0 += kMiniThreads)
                {
                    uint32_t const q = c0 + lane;
                    bool const f = q < pe && flags[q];
                    uint64_t const ballot = __ballot(f);
                    if (f)
                        hashes[out + emitted + __builtin_popcountll(ballot & ((1ull << lane) - 1ull))] = vals[q];
                    emitted += static_cast<uint64_t>(__builtin_popcountll(ballot));
                }
            }
            else
            {
                uint32_t const incl = hibf_amd::wave_inclusive_add_u32(static_cast<uint32_t>(__builtin_popcount(mine)));
                emitted += static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(incl), 63));
            }
            live = x != kMiniSink;
            p = b0 + (live ? x : 0u);
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

// Checks a shape and window as hibf_gpu_minimiser_hash documents them and
// fills the kernel parameters; returns a status (HIBF_GPU_OK when valid).
int minimiser_params(uint64_t shape, uint32_t window_size, uint64_t seed, MiniParams & mp)
{
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
    mp = MiniParams{};
    mp.shape = shape;
    mp.seed = seed;
    mp.size = size;
    mp.window = static_cast<uint32_t>(w);
    mp.gapped = shape != (1ull << size) - 1ull ? 1u : 0u;
    mp.words_cap = (kMiniSeg + mp.window + size - 1u + 255u) / 256u * 8u + 1u;
    return HIBF_GPU_OK;
}

unsigned minimiser_blocks(uint64_t n)
{
    return static_cast<unsigned>(std::max<uint64_t>(1u, std::min<uint64_t>(n, 1u << 16)));
}

} // namespace

namespace hibf_amd
{

int minimiser_check(uint64_t shape, uint32_t window_size)
{
    MiniParams mp;
    return minimiser_params(shape, window_size, 0u, mp);
}

hipError_t minimiser_scan_bytes(uint64_t n, size_t & bytes)
{
    bytes = 0;
    return n ? inclusive_sum_u64(nullptr, bytes, nullptr, nullptr, n, nullptr) : hipSuccess;
}

// Minimisers of n reads, asynchronous on s: the count pass, the offsets scan in
// scan_tmp (minimiser_scan_bytes), the thresholds uint16_t(fraction * count)
// when d_thresholds is given, and the write pass.  d_hashes must hold every
// minimiser (at most one per base).  The shape and window must have passed
// minimiser_check.
hipError_t minimiser_hash_async(uint8_t const * d_seq,
                                uint64_t const * d_seq_offsets,
                                uint64_t n,
                                uint64_t shape,
                                uint32_t window_size,
                                uint64_t seed,
                                uint64_t * d_hash_offsets,
                                uint64_t * d_hashes,
                                void * scan_tmp,
                                size_t scan_bytes,
                                double fraction,
                                uint16_t * d_thresholds,
                                hipStream_t s)
{
    MiniParams mp;
    if (minimiser_params(shape, window_size, seed, mp) != HIBF_GPU_OK)
        return hipErrorInvalidValue;
    size_t const lds = minimiser_lds_bytes(mp);
    hipLaunchKernelGGL(zero_first_kernel, dim3(1), dim3(64), 0, s, d_hash_offsets);
    if (n == 0)
        return hipGetLastError();
    hipLaunchKernelGGL(minimiser_kernel<false>, dim3(minimiser_blocks(n)), dim3(kMiniThreads), lds, s, d_seq,
                       d_seq_offsets, n, mp, d_hash_offsets + 1, nullptr, nullptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = inclusive_sum_u64(scan_tmp, scan_bytes, d_hash_offsets + 1, d_hash_offsets + 1, n, s);
    if (e == hipSuccess && d_thresholds)
    {
        hipLaunchKernelGGL(kmer_threshold_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, d_hash_offsets, n,
                           fraction, d_thresholds);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
    {
        hipLaunchKernelGGL(minimiser_kernel<true>, dim3(minimiser_blocks(n)), dim3(kMiniThreads), lds, s, d_seq,
                           d_seq_offsets, n, mp, nullptr, d_hash_offsets, d_hashes);
        e = hipGetLastError();
    }
    return e;
}

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
    MiniParams mp;
    if (int const rc = minimiser_params(shape, window_size, seed, mp))
        return rc;
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
