// gather_ceiling: the random-row read rate of the device, the practical
// ceiling of the IBF row fetch (one random row per hash function and value).
//
// Reads rows of ROW bytes at uniformly random row indices of a buffer much
// larger than every cache (default 16 GiB) and reports rows/s and row GB/s.
// Each index is umul64hi(random, n_rows) < n_rows, so every load is in
// bounds.  Rows of up to 128 bytes cost one whole cache line each; their rate
// is the line rate, which bounds the HIBF descent (32-byte rows of 256-bin
// IBFs); 1 KiB rows bound the flat count kernel (8,192 bins).
//
// Usage: gather_ceiling [GiB]      one JSON line per row size
//        gather_ceiling --window      32-byte rows over buffers of 64 MiB to
//                                     16 GiB: the rate inside the caches
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>

namespace
{
void check(hipError_t e, int line)
{
    if (e != hipSuccess)
    {
        std::fprintf(stderr, "gather_ceiling: %s (line %d)\n", hipGetErrorString(e), line);
        std::exit(1);
    }
}
#define CHECK(x) check((x), __LINE__)

__device__ __forceinline__ uint64_t fmix(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    x ^= x >> 33;
    return x;
}

// ROW / 16 lanes read one row with 16-byte loads; a wave reads 64 * 16 / ROW
// rows per load instruction and keeps U of them in flight per lane.
template <int ROW, int U>
__global__ void __launch_bounds__(256) gather_kernel(uint4 const * __restrict__ buf,
                                                     uint64_t n_rows,
                                                     uint32_t rounds,
                                                     uint64_t seed,
                                                     uint64_t * __restrict__ sink)
{
    constexpr int lanes_per_row = ROW / 16;
    int const lane = static_cast<int>(threadIdx.x & 63u);
    uint64_t const wave = static_cast<uint64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    uint64_t const stream_id = wave * (64 / lanes_per_row) + lane / lanes_per_row;
    int const part = lane % lanes_per_row;
    uint64_t acc = 0;
    for (uint32_t r = 0; r < rounds; ++r)
    {
        uint4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
        {
            uint64_t const key = seed ^ (stream_id * 0x9E3779B97F4A7C15ULL) ^ (static_cast<uint64_t>(r * U + u) << 40);
            uint64_t const row = __umul64hi(fmix(key), n_rows);
            v[u] = buf[row * lanes_per_row + part];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            acc += static_cast<uint64_t>(v[u].x) + v[u].y + v[u].z + v[u].w;
    }
    sink[wave * 64u + static_cast<uint64_t>(lane)] = acc; // keeps every lane's loads live
}

template <int ROW, int U>
void measure(uint4 const * buf, uint64_t bytes, uint64_t * sink, unsigned blocks, uint32_t rounds)
{
    uint64_t const n_rows = bytes / ROW;
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    hipLaunchKernelGGL((gather_kernel<ROW, U>), dim3(blocks), dim3(256), 0, 0, buf, n_rows, rounds, 1, sink);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    int const reps = 5;
    CHECK(hipEventRecord(a));
    for (int rep = 0; rep < reps; ++rep)
        hipLaunchKernelGGL((gather_kernel<ROW, U>), dim3(blocks), dim3(256), 0, 0, buf, n_rows, rounds, rep + 2,
                           sink);
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    double const s = ms * 1e-3 / reps;
    double const rows = static_cast<double>(blocks) * 4.0 * (64 / (ROW / 16)) * rounds * U;
    std::printf("{\"row_bytes\": %d, \"in_flight\": %d, \"blocks\": %u, \"rows\": %.0f, \"ms\": %.4f, "
                "\"rows_per_s\": %.4e, \"row_GBs\": %.1f, \"line_GBs\": %.1f}\n",
                ROW, U, blocks, rows, s * 1e3, rows / s, rows * ROW / s / 1e9,
                rows * (ROW < 128 ? 128 : ROW) / s / 1e9);
    std::fflush(stdout);
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
}
} // namespace

int main(int argc, char ** argv)
{
    if (argc > 1 && std::string(argv[1]) == "--window")
    {
        uint64_t const max_bytes = 16ull << 30;
        uint4 * buf = nullptr;
        uint64_t * sink = nullptr;
        unsigned const blocks = 4096;
        CHECK(hipMalloc(&buf, max_bytes));
        CHECK(hipMalloc(&sink, static_cast<uint64_t>(blocks) * 256u * sizeof(uint64_t)));
        CHECK(hipMemset(buf, 0x5a, max_bytes));
        CHECK(hipDeviceSynchronize());
        for (uint64_t mib : {64ull, 128ull, 256ull, 512ull, 1024ull, 4096ull, 16384ull})
        {
            std::printf("{\"window_MiB\": %llu}\n", static_cast<unsigned long long>(mib));
            measure<32, 8>(buf, mib << 20, sink, blocks, 64);
        }
        CHECK(hipFree(buf));
        CHECK(hipFree(sink));
        return 0;
    }
    uint64_t const gib = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 16;
    if (gib < 1 || gib > 128)
    {
        std::fprintf(stderr, "gather_ceiling: buffer size must be 1..128 GiB\n");
        return 2;
    }
    uint64_t const bytes = gib << 30;
    unsigned const blocks = 4096;
    uint4 * buf = nullptr;
    uint64_t * sink = nullptr;
    CHECK(hipMalloc(&buf, bytes));
    CHECK(hipMalloc(&sink, static_cast<uint64_t>(blocks) * 256u * sizeof(uint64_t)));
    CHECK(hipMemset(buf, 0x5a, bytes));
    CHECK(hipDeviceSynchronize());
    measure<32, 8>(buf, bytes, sink, blocks, 64);
    measure<64, 8>(buf, bytes, sink, blocks, 64);
    measure<128, 8>(buf, bytes, sink, blocks, 64);
    measure<256, 8>(buf, bytes, sink, blocks, 32);
    measure<1024, 8>(buf, bytes, sink, blocks, 32);
    CHECK(hipFree(buf));
    CHECK(hipFree(sink));
    return 0;
}
