// Union estimates of contiguous runs of user bins on the device.
//
// The layout DP (src/layout/hierarchical_binning.cpp:113-171) asks, for every
// column j of a level's sorted user bins and every k < j, for the estimate of
// the union of the sketches k .. j (src/sketch/toolbox.cpp:35-51:
// merge_and_estimate down from j).  That is n^2 / 2 estimates of 2^bits
// registers each, the reference's dominant build cost on many user bins.
//
// Here 8 threads own one k, one per float lane of the estimate
// (src/sketch/hyperloglog.cpp:66-105: lane L adds, over the registers in
// order, the terms of registers 8 q + 7 - L).  A thread keeps its lane's
// registers of the running union of sketches k .. j in VGPRs and walks j
// upward: per step it loads its lane's slice of sketch j (the same addresses
// for the wave's 8 values of k), takes the byte-wise max, and adds the lane's
// 2^-register terms in the reference's order, so each estimate's float sums
// are the reference's bit for bit.  The group then adds its 8 lane sums in
// lane order, divides alpha m^2 by the sum in double and, below 2.5 m,
// applies the small-value correction from a table m ln(m / zeros) computed on
// the host with the reference's expression.  Sketches are stored lane-major
// so that a lane's registers are contiguous.
#include "hll_union.hpp"
#include "kernels.hpp"

#include <algorithm>
#include <cmath>
#include <mutex>
#include <stdexcept>
#include <string>

#include "hibf_gpu.h"

namespace hibf_amd
{
int set_error(int status, std::string msg);
}

namespace
{

// Byte-wise max of two words of registers (every register < 128): bit 7 of
// byte i of (a | 0x80..) - b is set iff a_i >= b_i, with no borrow between bytes.
__device__ __forceinline__ uint32_t register_max(uint32_t a, uint32_t b)
{
    uint32_t const ge = ((a | 0x80808080u) - b) & 0x80808080u;
    uint32_t const mask = ge | (ge - (ge >> 7));
    return (a & mask) | (b & ~mask);
}

// Registers equal to zero in a word (every register < 128).
__device__ __forceinline__ uint32_t zero_registers(uint32_t x)
{
    return 4u - static_cast<uint32_t>(__popc((x + 0x7F7F7F7Fu) & 0x80808080u));
}

// 2^-r as a float (r <= 60): exponent field 127 - r.
__device__ __forceinline__ float inverse_power(uint32_t r)
{
    return __uint_as_float((127u - r) << 23);
}

__global__ void __launch_bounds__(256) lane_major_kernel(uint8_t const * __restrict__ regs,
                                                         uint64_t n,
                                                         uint32_t bits,
                                                         uint8_t * __restrict__ lanes)
{
    uint64_t const m = 1ULL << bits;
    uint64_t const total = n << bits;
    uint64_t const per_lane = m / 8u;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    {
        uint64_t const u = i >> bits;
        uint64_t const r = i & (m - 1u);
        uint64_t const L = r / per_lane;
        uint64_t const q = r % per_lane;
        lanes[i] = regs[(u << bits) + 8u * q + 7u - L];
    }
}

// Loads a lane slice of Q words (16-byte loads when Q is a multiple of 4:
// slices are then 16-byte aligned).
template <int Q>
__device__ __forceinline__ void load_slice(uint32_t const * src, uint32_t (&x)[Q])
{
    if constexpr (Q % 4 == 0)
    {
#pragma unroll
        for (int w = 0; w < Q; w += 4)
        {
            uint4 const v = reinterpret_cast<uint4 const *>(src)[w / 4];
            x[w] = v.x;
            x[w + 1] = v.y;
            x[w + 2] = v.z;
            x[w + 3] = v.w;
        }
    }
    else
    {
#pragma unroll
        for (int w = 0; w < Q; ++w)
            x[w] = src[w];
    }
}

// Q = 2^bits / 32 words per lane.  Thread t: k = t / 8, lane L = t % 8.  The
// wave walks j from its smallest k + 1 to np - 1; a thread takes part in step
// j when k < j.  Every thread runs the shuffles.
template <int Q>
__global__ void __launch_bounds__(256) union_triangle_kernel(uint8_t const * __restrict__ lanes,
                                                             uint32_t bits,
                                                             uint64_t const * __restrict__ pos,
                                                             uint64_t np,
                                                             double norm,
                                                             double const * __restrict__ small_estimate,
                                                             uint64_t * __restrict__ out)
{
    uint64_t const t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t const k = t >> 3;
    uint32_t const L = static_cast<uint32_t>(t & 7u);
    int const lane = static_cast<int>(threadIdx.x & 63u);
    int const first = lane & ~7;
    uint64_t const m = 1ULL << bits;
    uint64_t const wave_k = (t & ~63ULL) >> 3;
    if (wave_k + 1u >= np)
        return; // the whole wave has no column
    uint64_t const lane_bytes = m / 8u;

    uint32_t M[Q];
    {
        uint64_t const pk = pos[k < np ? k : np - 1u];
        load_slice<Q>(reinterpret_cast<uint32_t const *>(lanes + (pk << bits) + L * lane_bytes), M);
    }
    for (uint64_t j = wave_k + 1u; j < np; ++j)
    {
        bool const active = k < j;
        uint32_t const * src = reinterpret_cast<uint32_t const *>(lanes + (pos[j] << bits) + L * lane_bytes);
        float acc = 0.0f;
        if (active)
        {
            uint32_t S[Q];
            load_slice<Q>(src, S);
#pragma unroll
            for (int w = 0; w < Q; ++w)
                M[w] = register_max(M[w], S[w]);
#pragma unroll
            for (int w = 0; w < Q; ++w)
            {
                uint32_t const x = M[w];
                acc = acc + inverse_power(x & 0xFFu);
                acc = acc + inverse_power((x >> 8) & 0xFFu);
                acc = acc + inverse_power((x >> 16) & 0xFFu);
                acc = acc + inverse_power(x >> 24);
            }
        }
        float sum = 0.0f;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            sum = sum + __shfl(acc, first + i, 64);
        double e = norm / static_cast<double>(sum);
        // The zero registers matter only below 2.5 m (small unions): counted then.
        bool const small = active && e <= 2.5 * static_cast<double>(m);
        if (__any(small))
        {
            uint32_t zeros = 0;
            if (active)
#pragma unroll
                for (int w = 0; w < Q; ++w)
                    zeros += zero_registers(M[w]);
            uint32_t zs = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i)
                zs += static_cast<uint32_t>(__shfl(static_cast<int>(zeros), first + i, 64));
            if (small && zs != 0u)
                e = small_estimate[zs];
        }
        if (active && L == 0u)
            out[j * (j - 1u) / 2u + k] = static_cast<uint64_t>(e);
    }
}

// out[k] = estimate(merge(slot a, slot idx[k])) as a double, k < n: 8 threads
// per pair, thread L summing lane L of the merged registers in order.
template <int Q>
__global__ void __launch_bounds__(256) pair_union_kernel(uint8_t const * __restrict__ slots,
                                                         uint32_t bits,
                                                         uint32_t a,
                                                         uint32_t const * __restrict__ idx,
                                                         uint64_t n,
                                                         double norm,
                                                         double const * __restrict__ small_estimate,
                                                         double * __restrict__ out)
{
    uint64_t const t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t const k = t >> 3;
    uint32_t const L = static_cast<uint32_t>(t & 7u);
    int const first = static_cast<int>(threadIdx.x & 63u) & ~7;
    uint64_t const m = 1ULL << bits;
    uint64_t const lane_bytes = m / 8u;
    float acc = 0.0f;
    uint32_t const * const sa =
        reinterpret_cast<uint32_t const *>(slots + (static_cast<uint64_t>(a) << bits) + L * lane_bytes);
    uint32_t const * const sb =
        reinterpret_cast<uint32_t const *>(slots + (static_cast<uint64_t>(k < n ? idx[k] : a) << bits) + L * lane_bytes);
    if (k < n)
    {
        uint32_t A[Q], B[Q];
        load_slice<Q>(sa, A);
        load_slice<Q>(sb, B);
#pragma unroll
        for (int w = 0; w < Q; ++w)
        {
            uint32_t const x = register_max(A[w], B[w]);
            acc = acc + inverse_power(x & 0xFFu);
            acc = acc + inverse_power((x >> 8) & 0xFFu);
            acc = acc + inverse_power((x >> 16) & 0xFFu);
            acc = acc + inverse_power(x >> 24);
        }
    }
    float sum = 0.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        sum = sum + __shfl(acc, first + i, 64);
    double e = norm / static_cast<double>(sum);
    bool const small = k < n && e <= 2.5 * static_cast<double>(m);
    if (__any(small))
    {
        uint32_t zeros = 0;
        if (k < n)
        {
            uint32_t A[Q], B[Q];
            load_slice<Q>(sa, A);
            load_slice<Q>(sb, B);
#pragma unroll
            for (int w = 0; w < Q; ++w)
                zeros += zero_registers(register_max(A[w], B[w]));
        }
        uint32_t zs = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i)
            zs += static_cast<uint32_t>(__shfl(static_cast<int>(zeros), first + i, 64));
        if (small && zs != 0u)
            e = small_estimate[zs];
    }
    if (k < n && L == 0u)
        out[k] = e;
}

// Every pair i < j of the slots base .. base + L - 1: out[i L - i (i + 1) / 2 +
// j - i - 1] = estimate(merge(i, j)).  A block per row i (grid-stride), 32
// pairs per block pass, 8 threads per pair.
template <int Q>
__global__ void __launch_bounds__(256) all_pairs_kernel(uint8_t const * __restrict__ slots,
                                                        uint32_t bits,
                                                        uint64_t base,
                                                        uint64_t L,
                                                        double norm,
                                                        double const * __restrict__ small_estimate,
                                                        double * __restrict__ out)
{
    uint32_t const lane_id = static_cast<uint32_t>(threadIdx.x & 7u);
    int const first = static_cast<int>(threadIdx.x & 63u) & ~7;
    uint64_t const m = 1ULL << bits;
    uint64_t const lane_bytes = m / 8u;
    for (uint64_t i = blockIdx.x; i + 1u < L; i += gridDim.x)
    {
        uint32_t const * const sa = reinterpret_cast<uint32_t const *>(slots + ((base + i) << bits) + lane_id * lane_bytes);
        double * const row = out + (i * L - i * (i + 1u) / 2u);
        // Every thread of the block runs the same number of passes (the shuffles need the whole wave).
        for (uint64_t j0 = i + 1u; j0 < L; j0 += blockDim.x / 8u)
        {
            uint64_t const j = j0 + threadIdx.x / 8u;
            bool const valid = j < L;
            uint32_t const * const sb =
                reinterpret_cast<uint32_t const *>(slots + ((base + (valid ? j : i)) << bits) + lane_id * lane_bytes);
            float acc = 0.0f;
            if (valid)
            {
                uint32_t A[Q], B[Q];
                load_slice<Q>(sa, A);
                load_slice<Q>(sb, B);
#pragma unroll
                for (int w = 0; w < Q; ++w)
                {
                    uint32_t const x = register_max(A[w], B[w]);
                    acc = acc + inverse_power(x & 0xFFu);
                    acc = acc + inverse_power((x >> 8) & 0xFFu);
                    acc = acc + inverse_power((x >> 16) & 0xFFu);
                    acc = acc + inverse_power(x >> 24);
                }
            }
            float sum = 0.0f;
#pragma unroll
            for (int t = 0; t < 8; ++t)
                sum = sum + __shfl(acc, first + t, 64);
            double e = norm / static_cast<double>(sum);
            bool const small = valid && e <= 2.5 * static_cast<double>(m);
            if (__any(small))
            {
                uint32_t zeros = 0;
                if (valid)
                {
                    uint32_t A[Q], B[Q];
                    load_slice<Q>(sa, A);
                    load_slice<Q>(sb, B);
#pragma unroll
                    for (int w = 0; w < Q; ++w)
                        zeros += zero_registers(register_max(A[w], B[w]));
                }
                uint32_t zs = 0;
#pragma unroll
                for (int t = 0; t < 8; ++t)
                    zs += static_cast<uint32_t>(__shfl(static_cast<int>(zeros), first + t, 64));
                if (small && zs != 0u)
                    e = small_estimate[zs];
            }
            if (valid && lane_id == 0u)
                row[j - i - 1u] = e;
        }
    }
}

// out[j] = the first value of column j (k = 0): the prefix unions.
__global__ void __launch_bounds__(256) column_heads_kernel(uint64_t const * __restrict__ tri,
                                                           uint64_t np,
                                                           uint64_t * __restrict__ out)
{
    uint64_t const j = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= 1u && j < np)
        out[j] = tri[j * (j - 1u) / 2u];
}

template <int Q>
hipError_t launch_triangle(uint8_t const * lanes, uint32_t bits, uint64_t const * pos, uint64_t np, double norm,
                           double const * table, uint64_t * out, hipStream_t s)
{
    uint64_t const threads = (np - 1u) * 8u;
    unsigned const blocks = static_cast<unsigned>((threads + 255u) / 256u);
    hipLaunchKernelGGL(union_triangle_kernel<Q>, dim3(blocks), dim3(256), 0, s, lanes, bits, pos, np, norm, table, out);
    return hipGetLastError();
}

template <int Q>
hipError_t launch_pairs(uint8_t const * slots, uint32_t bits, uint32_t a, uint32_t const * idx, uint64_t n, double norm,
                        double const * table, double * out, hipStream_t s)
{
    unsigned const blocks = static_cast<unsigned>((n * 8u + 255u) / 256u);
    hipLaunchKernelGGL(pair_union_kernel<Q>, dim3(blocks), dim3(256), 0, s, slots, bits, a, idx, n, norm, table, out);
    return hipGetLastError();
}

hipError_t pair_unions(uint8_t const * slots, uint32_t bits, uint32_t a, uint32_t const * idx, uint64_t n, double norm,
                       double const * table, double * out, hipStream_t s)
{
    switch (bits)
    {
    case 5: return launch_pairs<1>(slots, bits, a, idx, n, norm, table, out, s);
    case 6: return launch_pairs<2>(slots, bits, a, idx, n, norm, table, out, s);
    case 7: return launch_pairs<4>(slots, bits, a, idx, n, norm, table, out, s);
    case 8: return launch_pairs<8>(slots, bits, a, idx, n, norm, table, out, s);
    case 9: return launch_pairs<16>(slots, bits, a, idx, n, norm, table, out, s);
    case 10: return launch_pairs<32>(slots, bits, a, idx, n, norm, table, out, s);
    case 11: return launch_pairs<64>(slots, bits, a, idx, n, norm, table, out, s);
    default: return launch_pairs<128>(slots, bits, a, idx, n, norm, table, out, s);
    }
}

template <int Q>
hipError_t launch_all_pairs(uint8_t const * slots, uint32_t bits, uint64_t base, uint64_t L, double norm,
                            double const * table, double * out, hipStream_t s)
{
    unsigned const blocks = static_cast<unsigned>(std::min<uint64_t>(L, 1u << 16));
    hipLaunchKernelGGL(all_pairs_kernel<Q>, dim3(blocks), dim3(256), 0, s, slots, bits, base, L, norm, table, out);
    return hipGetLastError();
}

hipError_t all_pairs_launch(uint8_t const * slots, uint32_t bits, uint64_t base, uint64_t L, double norm,
                            double const * table, double * out, hipStream_t s)
{
    switch (bits)
    {
    case 5: return launch_all_pairs<1>(slots, bits, base, L, norm, table, out, s);
    case 6: return launch_all_pairs<2>(slots, bits, base, L, norm, table, out, s);
    case 7: return launch_all_pairs<4>(slots, bits, base, L, norm, table, out, s);
    case 8: return launch_all_pairs<8>(slots, bits, base, L, norm, table, out, s);
    case 9: return launch_all_pairs<16>(slots, bits, base, L, norm, table, out, s);
    case 10: return launch_all_pairs<32>(slots, bits, base, L, norm, table, out, s);
    case 11: return launch_all_pairs<64>(slots, bits, base, L, norm, table, out, s);
    default: return launch_all_pairs<128>(slots, bits, base, L, norm, table, out, s);
    }
}

double normalization(uint32_t bits)
{
    uint64_t const size = 1ULL << bits;
    double const alpha = size == 32u ? 0.697 : size == 64u ? 0.709 : 0.7213 / (1.0 + 1.079 / static_cast<double>(size));
    return alpha * static_cast<double>(size) * static_cast<double>(size);
}

// Per device and bit width: m ln(m / z) for z = 0 .. m (entry 0 unused),
// computed once with the reference's expression and kept on the device for the
// process.
double const * small_estimate_table(uint32_t bits, hipError_t & e)
{
    constexpr int kMaxDevices = 64;
    static double * tables[kMaxDevices][hibf_amd::kUnionMaxBits + 1] = {};
    static std::mutex lock;
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess)
        return nullptr;
    if (dev < 0 || dev >= kMaxDevices)
    {
        e = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> const guard{lock};
    if (tables[dev][bits])
        return tables[dev][bits];
    uint64_t const size = 1ULL << bits;
    std::vector<double> h(size + 1u, 0.0);
    for (uint64_t z = 1; z <= size; ++z)
        h[z] = static_cast<double>(size) * std::log(static_cast<double>(size) / static_cast<double>(z));
    double * d = nullptr;
    e = hibf_amd::device_malloc(&d, h.size() * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess)
    {
        if (d)
            (void)hipFree(d);
        return nullptr;
    }
    tables[dev][bits] = d;
    return d;
}

int device_fail(hipError_t e, char const * what)
{
    (void)hipGetLastError();
    return hibf_amd::set_error(HIBF_GPU_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

} // namespace

namespace hibf_amd
{

hipError_t hll_lane_major(uint8_t const * d_regs, uint64_t n, uint32_t bits, uint8_t * d_lanes, hipStream_t s)
{
    uint64_t const total = n << bits;
    if (total == 0)
        return hipSuccess;
    unsigned const blocks = static_cast<unsigned>(std::min<uint64_t>((total + 255u) / 256u, 1u << 20));
    hipLaunchKernelGGL(lane_major_kernel, dim3(blocks), dim3(256), 0, s, d_regs, n, bits, d_lanes);
    return hipGetLastError();
}

hipError_t hll_union_triangle(uint8_t const * d_lanes,
                              uint32_t bits,
                              uint64_t const * d_pos,
                              uint64_t np,
                              uint64_t * d_out,
                              hipStream_t s)
{
    if (np < 2u)
        return hipSuccess;
    if (bits < 5u || bits > kUnionMaxBits)
        return hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    double const * table = small_estimate_table(bits, e);
    if (!table)
        return e;
    double const norm = normalization(bits);
    switch (bits)
    {
    case 5: return launch_triangle<1>(d_lanes, bits, d_pos, np, norm, table, d_out, s);
    case 6: return launch_triangle<2>(d_lanes, bits, d_pos, np, norm, table, d_out, s);
    case 7: return launch_triangle<4>(d_lanes, bits, d_pos, np, norm, table, d_out, s);
    case 8: return launch_triangle<8>(d_lanes, bits, d_pos, np, norm, table, d_out, s);
    case 9: return launch_triangle<16>(d_lanes, bits, d_pos, np, norm, table, d_out, s);
    case 10: return launch_triangle<32>(d_lanes, bits, d_pos, np, norm, table, d_out, s);
    case 11: return launch_triangle<64>(d_lanes, bits, d_pos, np, norm, table, d_out, s);
    default: return launch_triangle<128>(d_lanes, bits, d_pos, np, norm, table, d_out, s);
    }
}

DeviceUnionEstimates::~DeviceUnionEstimates()
{
    for (void * p : {static_cast<void *>(d_lanes_), static_cast<void *>(d_pos_), static_cast<void *>(d_tri_)})
        if (p)
            (void)hipFree(p);
    if (h_window_)
        (void)hipHostFree(h_window_);
    if (e0_)
        (void)hipEventDestroy(e0_);
    if (e1_)
        (void)hipEventDestroy(e1_);
    if (s_)
        (void)hipStreamDestroy(s_);
}

int DeviceUnionEstimates::create(uint8_t const * d_registers,
                                 uint64_t n,
                                 uint32_t bits,
                                 std::unique_ptr<DeviceUnionEstimates> & out)
{
    if (bits < 5u || bits > kUnionMaxBits)
        return set_error(HIBF_GPU_ERR_UNSUPPORTED, "device union estimates take sketches of 5 to 12 bits");
    std::unique_ptr<DeviceUnionEstimates> u{new DeviceUnionEstimates{}};
    u->bits_ = bits;
    u->n_ = n;
    u->window_cap_ = 8u << 20; // 64 MiB of pinned window
    hipError_t e = hipStreamCreateWithFlags(&u->s_, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipEventCreate(&u->e0_);
    if (e == hipSuccess)
        e = hipEventCreate(&u->e1_);
    if (e == hipSuccess)
        e = hibf_amd::device_malloc(reinterpret_cast<void **>(&u->d_lanes_), std::max<uint64_t>(n << bits, 1u));
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void **>(&u->h_window_), u->window_cap_ * sizeof(uint64_t));
    if (e == hipSuccess)
        e = hll_lane_major(d_registers, n, bits, u->d_lanes_, u->s_);
    if (e == hipSuccess)
        e = hipStreamSynchronize(u->s_);
    if (e != hipSuccess)
        return device_fail(e, "union estimates setup");
    out = std::move(u);
    return HIBF_GPU_OK;
}

bool DeviceUnionEstimates::start(std::vector<uint64_t> const & positions)
{
    np_ = positions.size();
    window_j0_ = window_j1_ = 0;
    for (uint64_t const p : positions)
        if (p >= n_)
            return set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "user bin position out of range") == HIBF_GPU_OK;
    if (np_ < 2u)
        return true;
    uint64_t const values = np_ * (np_ - 1u) / 2u;
    hipError_t e = hipSuccess;
    if (np_ > pos_cap_)
    {
        if (d_pos_)
            (void)hipFree(d_pos_);
        d_pos_ = nullptr;
        e = hibf_amd::device_malloc(reinterpret_cast<void **>(&d_pos_), np_ * sizeof(uint64_t));
        pos_cap_ = e == hipSuccess ? np_ : 0u;
    }
    if (e == hipSuccess && values > tri_cap_)
    {
        if (d_tri_)
            (void)hipFree(d_tri_);
        d_tri_ = nullptr;
        e = hibf_amd::device_malloc(reinterpret_cast<void **>(&d_tri_), values * sizeof(uint64_t));
        tri_cap_ = e == hipSuccess ? values : 0u;
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_pos_, positions.data(), np_ * sizeof(uint64_t), hipMemcpyHostToDevice, s_);
    if (e == hipSuccess)
        e = hipEventRecord(e0_, s_);
    if (e == hipSuccess)
        e = hll_union_triangle(d_lanes_, bits_, d_pos_, np_, d_tri_, s_);
    if (e == hipSuccess)
        e = hipEventRecord(e1_, s_);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s_);
    float ms = 0.0f;
    if (e == hipSuccess)
        e = hipEventElapsedTime(&ms, e0_, e1_);
    if (e != hipSuccess)
        return device_fail(e, "union estimates") == HIBF_GPU_OK;
    kernel_seconds += ms / 1e3;
    estimates += values;
    return true;
}

bool DeviceUnionEstimates::prefix(uint64_t * prefix)
{
    // The prefix unions are the k = 0 entries, the first value of each column,
    // gathered on the device.
    if (np_ < 2u)
        return true;
    uint64_t * d_heads = nullptr;
    hipError_t e = hibf_amd::device_malloc(reinterpret_cast<void **>(&d_heads), np_ * sizeof(uint64_t));
    if (e == hipSuccess)
    {
        hipLaunchKernelGGL(column_heads_kernel, dim3(static_cast<unsigned>((np_ + 255u) / 256u)), dim3(256), 0, s_,
                           d_tri_, np_, d_heads);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(prefix + 1, d_heads + 1, (np_ - 1u) * sizeof(uint64_t), hipMemcpyDeviceToHost, s_);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s_);
    if (d_heads)
        (void)hipFree(d_heads);
    if (e != hipSuccess)
        return device_fail(e, "union estimates prefix") == HIBF_GPU_OK;
    return true;
}

// Copies columns j, j + 1, ... (as many as fit the window) to the host.
bool DeviceUnionEstimates::fetch(uint64_t j)
{
    uint64_t const base = j * (j - 1u) / 2u;
    uint64_t j1 = j + 1u;
    while (j1 < np_ && j1 * (j1 + 1u) / 2u - base <= window_cap_)
        ++j1;
    uint64_t const values = j1 * (j1 - 1u) / 2u - base;
    if (values > window_cap_)
        return set_error(HIBF_GPU_ERR_UNSUPPORTED, "a layout column exceeds the union-estimate window") == HIBF_GPU_OK;
    hipError_t e = hipMemcpyAsync(h_window_, d_tri_ + base, values * sizeof(uint64_t), hipMemcpyDeviceToHost, s_);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s_);
    if (e != hipSuccess)
        return device_fail(e, "union estimates copy") == HIBF_GPU_OK;
    window_j0_ = j;
    window_j1_ = j1;
    return true;
}

uint64_t const * DeviceUnionEstimates::column(uint64_t j)
{
    if (j == 0 || j >= np_)
        return nullptr;
    if ((j < window_j0_ || j >= window_j1_) && !fetch(j))
        return nullptr;
    return h_window_ + (j * (j - 1u) / 2u - window_j0_ * (window_j0_ - 1u) / 2u);
}


DevicePairUnions::~DevicePairUnions()
{
    for (void * p : {static_cast<void *>(d_slots_), static_cast<void *>(d_idx_), static_cast<void *>(d_out_)})
        if (p)
            (void)hipFree(p);
    for (void * p : {static_cast<void *>(h_stage_), static_cast<void *>(h_idx_), static_cast<void *>(h_out_)})
        if (p)
            (void)hipHostFree(p);
    if (s_)
        (void)hipStreamDestroy(s_);
}

int DevicePairUnions::create(uint32_t bits, std::unique_ptr<DevicePairUnions> & out)
{
    if (bits < 5u || bits > kUnionMaxBits)
        return set_error(HIBF_GPU_ERR_UNSUPPORTED, "device union estimates take sketches of 5 to 12 bits");
    std::unique_ptr<DevicePairUnions> u{new DevicePairUnions{}};
    u->bits_ = bits;
    u->norm_ = normalization(bits);
    hipError_t e = hipSuccess;
    u->d_table_ = small_estimate_table(bits, e);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&u->s_, hipStreamNonBlocking);
    if (e != hipSuccess)
        return device_fail(e, "pair unions setup");
    out = std::move(u);
    return HIBF_GPU_OK;
}

void DevicePairUnions::check(hipError_t e, char const * what)
{
    if (e != hipSuccess)
    {
        device_fail(e, what);
        throw std::runtime_error(what);
    }
}

void DevicePairUnions::reset()
{
    slots_.clear();
    used_ = uploaded_ = 0;
}

// Room for pairs_needed pairs and slots_needed slots (device slots keep their
// contents when they grow).
void DevicePairUnions::ensure(uint64_t pairs_needed, uint64_t slots_needed)
{
    uint64_t const m = 1ULL << bits_;
    if (pairs_needed > pair_cap_)
    {
        uint64_t const cap = std::max<uint64_t>(pairs_needed, 2u * pair_cap_);
        for (void * p : {static_cast<void *>(d_idx_), static_cast<void *>(d_out_)})
            if (p)
                (void)hipFree(p);
        for (void * p : {static_cast<void *>(h_idx_), static_cast<void *>(h_out_)})
            if (p)
                (void)hipHostFree(p);
        d_idx_ = nullptr;
        d_out_ = nullptr;
        h_idx_ = nullptr;
        h_out_ = nullptr;
        pair_cap_ = 0;
        check(hibf_amd::device_malloc(reinterpret_cast<void **>(&d_idx_), cap * sizeof(uint32_t)), "pair unions buffers");
        check(hibf_amd::device_malloc(reinterpret_cast<void **>(&d_out_), cap * sizeof(double)), "pair unions buffers");
        check(hipHostMalloc(reinterpret_cast<void **>(&h_idx_), cap * sizeof(uint32_t)), "pair unions buffers");
        check(hipHostMalloc(reinterpret_cast<void **>(&h_out_), cap * sizeof(double)), "pair unions buffers");
        pair_cap_ = cap;
    }
    if (slots_needed > slot_cap_)
    {
        uint64_t const cap = std::max<uint64_t>(slots_needed, 2u * slot_cap_);
        uint8_t * fresh = nullptr;
        check(hibf_amd::device_malloc(reinterpret_cast<void **>(&fresh), cap << bits_), "pair unions slots");
        if (d_slots_)
        {
            hipError_t const e = hipMemcpyAsync(fresh, d_slots_, uploaded_ << bits_, hipMemcpyDeviceToDevice, s_);
            if (e == hipSuccess)
                check(hipStreamSynchronize(s_), "pair unions slots");
            (void)hipFree(d_slots_);
            check(e, "pair unions slots");
        }
        d_slots_ = fresh;
        slot_cap_ = cap;
    }
    uint64_t const staged = slots_needed - uploaded_;
    if (staged > stage_cap_)
    {
        uint64_t const cap = std::max<uint64_t>(staged, 2u * stage_cap_);
        if (h_stage_)
            (void)hipHostFree(h_stage_);
        h_stage_ = nullptr;
        stage_cap_ = 0;
        check(hipHostMalloc(reinterpret_cast<void **>(&h_stage_), cap * m), "pair unions staging");
        stage_cap_ = cap;
    }
}

uint32_t DevicePairUnions::slot_of(seqan::hibf::sketch::hyperloglog const * h)
{
    auto const [it, inserted] = slots_.try_emplace(h, static_cast<uint32_t>(used_));
    if (inserted)
    {
        pending_.push_back(h);
        ++used_;
    }
    return it->second;
}

// Lane-major copies of the sketches first seen in this call (lane L's
// registers 8 q + 7 - L), queued to their slots.
void DevicePairUnions::upload_pending()
{
    uint64_t const m = 1ULL << bits_;
    for (uint64_t p = 0; p < pending_.size(); ++p)
    {
        std::vector<uint8_t> const & r = pending_[p]->registers();
        if (r.size() != m)
        {
            set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "sketches of different bit widths");
            throw std::runtime_error("sketch width");
        }
        uint8_t * const dst = h_stage_ + p * m;
        for (uint64_t L = 0; L < 8u; ++L)
            for (uint64_t q = 0; q < m / 8u; ++q)
                dst[L * (m / 8u) + q] = r[8u * q + 7u - L];
    }
    if (!pending_.empty())
        check(hipMemcpyAsync(d_slots_ + (uploaded_ << bits_), h_stage_, pending_.size() * m, hipMemcpyHostToDevice, s_),
              "pair unions upload");
}

void DevicePairUnions::all_pairs(std::vector<seqan::hibf::sketch::hyperloglog const *> const & leaves, double * out)
{
    uint64_t const L = leaves.size();
    if (L < 2u)
        return;
    // The leaves take consecutive slots (a clustering starts with them).
    pending_.clear();
    uint64_t const base = used_;
    bool consecutive = true;
    for (uint64_t i = 0; i < L && consecutive; ++i)
        consecutive = slot_of(leaves[i]) == base + i;
    ensure(0, used_);
    upload_pending();
    uploaded_ = used_;
    if (!consecutive) // a leaf seen before in this clustering: row by row
        return union_source::all_pairs(leaves, out);
    uint64_t const pairs_n = L * (L - 1u) / 2u;
    double * d_tri = nullptr;
    check(hibf_amd::device_malloc(reinterpret_cast<void **>(&d_tri), pairs_n * sizeof(double)), "pair unions triangle");
    hipError_t e = all_pairs_launch(d_slots_, bits_, base, L, norm_, d_table_, d_tri, s_);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d_tri, pairs_n * sizeof(double), hipMemcpyDeviceToHost, s_);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s_);
    (void)hipFree(d_tri);
    check(e, "pair unions triangle");
    ++calls;
    pairs += pairs_n;
}

void DevicePairUnions::unions(seqan::hibf::sketch::hyperloglog const & a,
                              std::vector<seqan::hibf::sketch::hyperloglog const *> const & others,
                              double * out)
{
    uint64_t const n = others.size();
    if (n == 0)
        return;
    uint64_t const m = 1ULL << bits_;
    pending_.clear();
    uint32_t const a_slot = slot_of(&a);
    std::vector<uint32_t> idx(n);
    for (uint64_t k = 0; k < n; ++k)
        idx[k] = slot_of(others[k]);
    ensure(n, used_);
    upload_pending();
    std::copy(idx.begin(), idx.end(), h_idx_);
    check(hipMemcpyAsync(d_idx_, h_idx_, n * sizeof(uint32_t), hipMemcpyHostToDevice, s_), "pair unions upload");
    check(pair_unions(d_slots_, bits_, a_slot, d_idx_, n, norm_, d_table_, d_out_, s_), "pair unions");
    check(hipMemcpyAsync(h_out_, d_out_, n * sizeof(double), hipMemcpyDeviceToHost, s_), "pair unions copy");
    check(hipStreamSynchronize(s_), "pair unions");
    uploaded_ = used_;
    std::copy(h_out_, h_out_ + n, out);
    ++calls;
    pairs += n;
}

} // namespace hibf_amd

extern "C" {

int hibf_gpu_hll_union_estimates(uint8_t const * d_registers,
                                 uint64_t n_sketches,
                                 uint32_t bits,
                                 uint64_t const * d_positions,
                                 uint64_t n_positions,
                                 uint64_t * d_out,
                                 void * stream)
{
    if (bits < 5u || bits > hibf_amd::kUnionMaxBits)
        return hibf_amd::set_error(HIBF_GPU_ERR_UNSUPPORTED, "device union estimates take sketches of 5 to 12 bits");
    if (n_positions < 2u)
        return HIBF_GPU_OK;
    if (!d_registers || !d_positions || !d_out)
        return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    std::vector<uint64_t> pos(n_positions);
    hipError_t e = hipMemcpyAsync(pos.data(), d_positions, n_positions * sizeof(uint64_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return device_fail(e, "union estimates positions");
    for (uint64_t const p : pos)
        if (p >= n_sketches)
            return hibf_amd::set_error(HIBF_GPU_ERR_INVALID_ARGUMENT, "position out of range");
    uint8_t * lanes = nullptr;
    e = hibf_amd::device_malloc(reinterpret_cast<void **>(&lanes), n_sketches << bits);
    if (e == hipSuccess)
        e = hibf_amd::hll_lane_major(d_registers, n_sketches, bits, lanes, s);
    if (e == hipSuccess)
        e = hibf_amd::hll_union_triangle(lanes, bits, d_positions, n_positions, d_out, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (lanes)
        (void)hipFree(lanes);
    return e == hipSuccess ? HIBF_GPU_OK : device_fail(e, "union estimates");
}

} // extern "C"
