// HyperLogLog sketches on the device: the build's user-bin cardinality
// estimates (src/sketch/compute_sketches.cpp:27-43 fills one sketch per user
// bin, estimate_kmer_counts.hpp truncates each estimate to size_t), and the
// register-wise merge the layout's union estimates use.
//
// A sketch is 2^bits uint8 registers (src/sketch/hyperloglog.cpp:26-46).  For
// a value v: h = lo64(v * C) ^ hi64(v * C) with C = 0x9E3779B97F4A7C15
// (hyperloglog.cpp:48-54); register h >> (64 - bits) keeps the max of
// countl_zero((h << bits) | (2^bits - 1)) + 1 (:56-64).
//
// Sketching: the value sets are cut into chunks of at most 2^16 values; a
// block builds a chunk's registers in LDS (4 registers per 32-bit word, byte
// max by compare-and-swap, which is rare once the registers fill up) and
// writes them out, with a byte-max compare-and-swap on the HBM words when the
// sketch has several chunks.  Sketches wider than 2^16 registers skip LDS.
// Estimating: 8 threads per sketch add 2^-register in the reference's eight
// float lanes and lane order (:66-105); the division and the small-value
// correction (:107-126) run on the host in double, as the reference does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "device_common.hpp"
#include "hibf_gpu.h"
#include "kernels.hpp"

namespace hibf_amd
{
int set_error(int status, std::string msg);
}

namespace
{

using hibf_amd::synth_value;

constexpr uint64_t kHllChunk = 1u << 16;   // values per work item
constexpr uint32_t kHllLdsMaxBits = 16u;   // 64 KiB of LDS registers

struct HllChunk
{
    uint64_t sketch;
    uint64_t begin; // values index (CSR) or index within the user bin (synthetic)
    uint64_t end;
    uint64_t shared; // the sketch has more than one chunk
};

__device__ __forceinline__ uint64_t hll_hash(uint64_t v)
{
    uint64_t const c = 0x9E3779B97F4A7C15ULL;
    return (v * c) ^ __umul64hi(v, c);
}

__device__ __forceinline__ uint32_t bytes_max(uint32_t a, uint32_t b)
{
    uint32_t r = 0;
#pragma unroll
    for (int s = 0; s < 32; s += 8)
    {
        uint32_t const x = (a >> s) & 0xFFu;
        uint32_t const y = (b >> s) & 0xFFu;
        r |= (x > y ? x : y) << s;
    }
    return r;
}

// Byte `idx` of the word array p becomes max(byte, rank).
template <typename Word>
__device__ __forceinline__ void byte_max(Word * p, uint64_t idx, uint32_t rank)
{
    uint32_t * const w = reinterpret_cast<uint32_t *>(p) + (idx >> 2);
    uint32_t const shift = static_cast<uint32_t>(idx & 3u) * 8u;
    uint32_t old = __atomic_load_n(w, __ATOMIC_RELAXED);
    while (((old >> shift) & 0xFFu) < rank)
    {
        uint32_t const want = (old & ~(0xFFu << shift)) | (rank << shift);
        uint32_t const seen = atomicCAS(w, old, want);
        if (seen == old)
            break;
        old = seen;
    }
}

// Word w becomes the byte-wise max of itself and x.
__device__ __forceinline__ void word_max(uint32_t * w, uint32_t x)
{
    uint32_t old = __atomic_load_n(w, __ATOMIC_RELAXED);
    while (true)
    {
        uint32_t const want = bytes_max(old, x);
        if (want == old)
            break;
        uint32_t const seen = atomicCAS(w, old, want);
        if (seen == old)
            break;
        old = seen;
    }
}

__device__ __forceinline__ void hll_index_rank(uint64_t v, uint32_t bits, uint64_t & idx, uint32_t & rank)
{
    uint64_t const h = hll_hash(v);
    idx = h >> (64u - bits);
    uint64_t const rest = (h << bits) | ((1ULL << bits) - 1u);
    rank = static_cast<uint32_t>(__clzll(static_cast<long long>(rest))) + 1u;
}

__global__ void __launch_bounds__(256) fill_zero_kernel(uint4 * p, uint64_t n16)
{
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n16;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        p[i] = make_uint4(0u, 0u, 0u, 0u);
}

template <bool SYNTH>
__global__ void __launch_bounds__(256) hll_lds_kernel(HllChunk const * __restrict__ chunks,
                                                      uint64_t n_chunks,
                                                      uint64_t const * __restrict__ values,
                                                      uint64_t seed,
                                                      uint32_t bits,
                                                      uint8_t * __restrict__ regs)
{
    extern __shared__ uint32_t lds[]; // 2^bits registers, 4 per word
    uint32_t const words = (1u << bits) / 4u;
    int const tid = static_cast<int>(threadIdx.x);
    int const nt = static_cast<int>(blockDim.x);
    for (uint64_t c = blockIdx.x; c < n_chunks; c += gridDim.x)
    {
        HllChunk const ch = chunks[c];
        for (uint32_t i = static_cast<uint32_t>(tid); i < words; i += static_cast<uint32_t>(nt))
            lds[i] = 0u;
        __syncthreads();
        for (uint64_t k = ch.begin + static_cast<uint64_t>(tid); k < ch.end; k += static_cast<uint64_t>(nt))
        {
            uint64_t const v = SYNTH ? synth_value(seed, ch.sketch, k) : values[k];
            uint64_t idx;
            uint32_t rank;
            hll_index_rank(v, bits, idx, rank);
            byte_max(lds, idx, rank);
        }
        __syncthreads();
        uint32_t * const out = reinterpret_cast<uint32_t *>(regs + (ch.sketch << bits));
        for (uint32_t i = static_cast<uint32_t>(tid); i < words; i += static_cast<uint32_t>(nt))
        {
            uint32_t const w = lds[i];
            if (!ch.shared)
                out[i] = w;
            else if (w)
                word_max(out + i, w);
        }
        __syncthreads();
    }
}

template <bool SYNTH>
__global__ void __launch_bounds__(256) hll_global_kernel(HllChunk const * __restrict__ chunks,
                                                         uint64_t n_chunks,
                                                         uint64_t const * __restrict__ values,
                                                         uint64_t seed,
                                                         uint32_t bits,
                                                         uint8_t * __restrict__ regs)
{
    for (uint64_t c = blockIdx.x; c < n_chunks; c += gridDim.x)
    {
        HllChunk const ch = chunks[c];
        uint8_t * const r = regs + (ch.sketch << bits);
        for (uint64_t k = ch.begin + threadIdx.x; k < ch.end; k += blockDim.x)
        {
            uint64_t const v = SYNTH ? synth_value(seed, ch.sketch, k) : values[k];
            uint64_t idx;
            uint32_t rank;
            hll_index_rank(v, bits, idx, rank);
            byte_max(r, idx, rank);
        }
    }
}

// 8 threads per sketch.  Thread L adds, for every 32-register block and
// j = 0, 8, 16, 24, the term of register j + 7 - L (lane L of the reference's
// _mm256_set_ps); the sketch's first thread then adds the 8 lanes in order.
__global__ void __launch_bounds__(256) hll_sum_kernel(uint8_t const * __restrict__ regs,
                                                      uint64_t n,
                                                      uint32_t bits,
                                                      float * __restrict__ sums,
                                                      uint32_t * __restrict__ zeros)
{
    uint64_t const t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t const s = t >> 3;
    uint32_t const L = static_cast<uint32_t>(t & 7u);
    int const lane = static_cast<int>(threadIdx.x & 63u);
    float part = 0.0f;
    uint32_t z = 0;
    if (s < n)
    {
        uint8_t const * const cur = regs + (s << bits);
        uint64_t const blocks = (1ULL << bits) / 32u;
        for (uint64_t b = 0; b < blocks; ++b)
            for (uint32_t j = 0; j <= 24u; j += 8u)
            {
                uint32_t const r = cur[32u * b + j + 7u - L];
                part = part + __builtin_ldexpf(1.0f, -static_cast<int>(r)); // 2^-r, exact
                z += r == 0u ? 1u : 0u;
            }
    }
    float sum = 0.0f;
    uint32_t zs = 0;
    int const first = lane & ~7;
    for (int k = 0; k < 8; ++k)
    {
        sum = sum + __shfl(part, first + k, 64);
        zs += static_cast<uint32_t>(__shfl(static_cast<int>(z), first + k, 64));
    }
    if (s < n && L == 0u)
    {
        sums[s] = sum;
        zeros[s] = zs;
    }
}

__global__ void __launch_bounds__(256) hll_merge_kernel(uint8_t const * __restrict__ regs,
                                                        uint32_t bits,
                                                        uint64_t const * __restrict__ members,
                                                        uint64_t const * __restrict__ group_offsets,
                                                        uint64_t n_groups,
                                                        uint8_t * __restrict__ out)
{
    uint64_t const words = (1ULL << bits) / 4u;
    for (uint64_t g = blockIdx.x; g < n_groups; g += gridDim.x)
    {
        uint32_t * const o = reinterpret_cast<uint32_t *>(out + (g << bits));
        uint64_t const m0 = group_offsets[g], m1 = group_offsets[g + 1];
        for (uint64_t w = threadIdx.x; w < words; w += blockDim.x)
        {
            uint32_t m = 0;
            for (uint64_t k = m0; k < m1; ++k)
                m = bytes_max(m, reinterpret_cast<uint32_t const *>(regs + (members[k] << bits))[w]);
            o[w] = m;
        }
    }
}

int device_error(hipError_t e, char const * what)
{
    (void)hipGetLastError();
    return hibf_amd::set_error(HIBF_GPU_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

int check_bits(uint32_t bits)
{
    if (bits < 5u || bits > 32u)
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "[HyperLogLog] bit width must be in the range [5,32].");
    return HIBF_GPU_OK;
}

// n sketches of 2^bits registers must be addressable in 64 bits.
int check_size(uint64_t n, uint32_t bits)
{
    if (n > (~0ULL >> bits))
        return hibf_amd::set_error(HIBF_GPU_ERR_UNSUPPORTED, "number of sketches x 2^bits registers exceeds 2^64 bytes");
    return HIBF_GPU_OK;
}

unsigned grid_for(uint64_t n, uint64_t cap)
{
    return static_cast<unsigned>(std::max<uint64_t>(1u, std::min<uint64_t>(n, cap)));
}

// Zeroes the registers, then runs the chunked sketch kernels.
int run_sketches(std::vector<HllChunk> const & chunks,
                 uint64_t n_sketches,
                 uint64_t const * d_values,
                 uint64_t seed,
                 bool synthetic,
                 uint32_t bits,
                 uint8_t * d_registers,
                 hipStream_t s)
{
    uint64_t const bytes = n_sketches << bits;
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid_for((bytes / 16u + 255u) / 256u, 1u << 16)), dim3(256), 0, s,
                       reinterpret_cast<uint4 *>(d_registers), bytes / 16u);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return device_error(e, "hll zero");
    if (chunks.empty())
    {
        e = hipStreamSynchronize(s);
        return e == hipSuccess ? HIBF_GPU_OK : device_error(e, "hll sketch");
    }
    void * d_chunks = nullptr;
    e = hibf_amd::device_malloc(&d_chunks, chunks.size() * sizeof(HllChunk));
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_chunks, chunks.data(), chunks.size() * sizeof(HllChunk), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
    {
        auto const * ch = static_cast<HllChunk const *>(d_chunks);
        unsigned const grid = grid_for(chunks.size(), 4096u);
        if (bits <= kHllLdsMaxBits)
        {
            size_t const lds = std::max<size_t>((1u << bits), 64u);
            if (synthetic)
                hipLaunchKernelGGL(hll_lds_kernel<true>, dim3(grid), dim3(256), lds, s, ch, chunks.size(), d_values,
                                   seed, bits, d_registers);
            else
                hipLaunchKernelGGL(hll_lds_kernel<false>, dim3(grid), dim3(256), lds, s, ch, chunks.size(), d_values,
                                   seed, bits, d_registers);
        }
        else if (synthetic)
            hipLaunchKernelGGL(hll_global_kernel<true>, dim3(grid), dim3(256), 0, s, ch, chunks.size(), d_values, seed,
                               bits, d_registers);
        else
            hipLaunchKernelGGL(hll_global_kernel<false>, dim3(grid), dim3(256), 0, s, ch, chunks.size(), d_values, seed,
                               bits, d_registers);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (d_chunks)
        (void)hipFree(d_chunks);
    return e == hipSuccess ? HIBF_GPU_OK : device_error(e, "hll sketch");
}

// Cuts [0, size) of sketch `sk` (values from `base`) into chunks.
void add_chunks(std::vector<HllChunk> & out, uint64_t sk, uint64_t base, uint64_t size)
{
    uint64_t const shared = size > kHllChunk ? 1u : 0u;
    for (uint64_t b = 0; b < size; b += kHllChunk)
        out.push_back({sk, base + b, base + std::min(size, b + kHllChunk), shared});
}

} // namespace

namespace hibf_amd
{

// hyperloglog::estimate (src/sketch/hyperloglog.cpp:31-46, 107-126) from the
// float sum and the zero-register count.
double hll_finish(uint32_t bits, float sum, uint32_t zeros)
{
    uint64_t const size = 1ULL << bits;
    double const alpha = size == 32u ? 0.697 : size == 64u ? 0.709 : 0.7213 / (1.0 + 1.079 / static_cast<double>(size));
    double const norm = alpha * static_cast<double>(size) * static_cast<double>(size);
    double estimate = norm / static_cast<double>(sum);
    if (estimate <= 2.5 * static_cast<double>(size) && zeros != 0u)
        estimate = static_cast<double>(size) * std::log(static_cast<double>(size) / zeros);
    return estimate;
}

// Estimates of n sketches resident on the device (synchronous).
int hll_estimates(uint8_t const * d_registers, uint64_t n, uint32_t bits, double * h_estimates, hipStream_t s)
{
    if (n == 0)
        return HIBF_GPU_OK;
    void * d = nullptr;
    hipError_t e = hibf_amd::device_malloc(&d, n * (sizeof(float) + sizeof(uint32_t)));
    std::vector<float> sums(n);
    std::vector<uint32_t> zeros(n);
    if (e == hipSuccess)
    {
        auto * d_sums = static_cast<float *>(d);
        auto * d_zeros = reinterpret_cast<uint32_t *>(d_sums + n);
        hipLaunchKernelGGL(hll_sum_kernel, dim3(static_cast<unsigned>((n * 8u + 255u) / 256u)), dim3(256), 0, s,
                           d_registers, n, bits, d_sums, d_zeros);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(sums.data(), d_sums, n * sizeof(float), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(zeros.data(), d_zeros, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipStreamSynchronize(s);
    }
    if (d)
        (void)hipFree(d);
    if (e != hipSuccess)
        return device_error(e, "hll estimate");
    for (uint64_t i = 0; i < n; ++i)
        h_estimates[i] = hll_finish(bits, sums[i], zeros[i]);
    return HIBF_GPU_OK;
}

// Sketches of the synthetic user bins (synth_value(seed, u, i), i < sizes[u]).
int hll_sketch_synthetic(uint64_t seed, uint64_t const * h_sizes, uint64_t n, uint32_t bits, uint8_t * d_registers,
                         hipStream_t s)
{
    std::vector<HllChunk> chunks;
    for (uint64_t u = 0; u < n; ++u)
        add_chunks(chunks, u, 0, h_sizes[u]);
    return run_sketches(chunks, n, nullptr, seed, true, bits, d_registers, s);
}

} // namespace hibf_amd

extern "C" {

int hibf_gpu_hll_sketch(uint64_t const * d_values,
                        uint64_t const * d_offsets,
                        uint64_t n_sketches,
                        uint32_t bits,
                        uint8_t * d_registers,
                        void * stream)
{
    if (int rc = check_bits(bits))
        return rc;
    if (int rc = check_size(n_sketches, bits))
        return rc;
    if ((!d_offsets || !d_registers) && n_sketches)
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (n_sketches == 0)
        return HIBF_GPU_OK;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    std::vector<uint64_t> offs(n_sketches + 1);
    hipError_t e = hipMemcpyAsync(offs.data(), d_offsets, offs.size() * 8u, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return device_error(e, "hll offsets");
    std::vector<HllChunk> chunks;
    for (uint64_t i = 0; i < n_sketches; ++i)
    {
        if (offs[i + 1] < offs[i])
            return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "offsets must not decrease");
        add_chunks(chunks, i, offs[i], offs[i + 1] - offs[i]);
    }
    if (!d_values && offs[n_sketches] != offs[0])
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null values");
    return run_sketches(chunks, n_sketches, d_values, 0, false, bits, d_registers, s);
}

int hibf_gpu_hll_sketch_synthetic(uint64_t seed,
                                  uint64_t const * h_sizes,
                                  uint64_t n_sketches,
                                  uint32_t bits,
                                  uint8_t * d_registers,
                                  void * stream)
{
    if (int rc = check_bits(bits))
        return rc;
    if (int rc = check_size(n_sketches, bits))
        return rc;
    if ((!h_sizes || !d_registers) && n_sketches)
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (n_sketches == 0)
        return HIBF_GPU_OK;
    return hibf_amd::hll_sketch_synthetic(seed, h_sizes, n_sketches, bits, d_registers,
                                          static_cast<hipStream_t>(stream));
}

int hibf_gpu_hll_merge(uint8_t const * d_registers,
                       uint32_t bits,
                       uint64_t const * d_members,
                       uint64_t const * d_group_offsets,
                       uint64_t n_groups,
                       uint8_t * d_out,
                       void * stream)
{
    if (int rc = check_bits(bits))
        return rc;
    if (int rc = check_size(n_groups, bits))
        return rc;
    if ((!d_registers || !d_members || !d_group_offsets || !d_out) && n_groups)
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (n_groups == 0)
        return HIBF_GPU_OK;
    hipLaunchKernelGGL(hll_merge_kernel, dim3(grid_for(n_groups, 1u << 16)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), d_registers, bits, d_members, d_group_offsets, n_groups, d_out);
    hipError_t const e = hipGetLastError();
    return e == hipSuccess ? HIBF_GPU_OK : device_error(e, "hll merge");
}

int hibf_gpu_hll_estimate(uint8_t const * d_registers, uint64_t n_sketches, uint32_t bits, double * h_estimates,
                          void * stream)
{
    if (int rc = check_bits(bits))
        return rc;
    if (int rc = check_size(n_sketches, bits))
        return rc;
    if ((!d_registers || !h_estimates) && n_sketches)
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    return hibf_amd::hll_estimates(d_registers, n_sketches, bits, h_estimates, static_cast<hipStream_t>(stream));
}

} // extern "C"
