// C-ABI of the MI355X-native HIBF query path (declared in include/hibf_gpu.h).
// Host-side validation, device memory management and kernel dispatch.  No
// exception crosses this boundary; every entry point returns a status code.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "hibf_gpu.h"
#include "build.hpp"
#include "kernels.hpp"
#include "hll_union.hpp"
#include "layout.hpp"

using namespace hibf_amd;

namespace
{

thread_local std::string g_last_error;

int fail(int status, std::string msg)
{
    g_last_error = std::move(msg);
    return status;
}

int hip_fail(hipError_t e, char const * what)
{
    (void)hipGetLastError(); // clear the sticky per-thread error
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
    if (e == hipErrorNoDevice)
        return fail(HIBF_GPU_ERR_NO_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    return fail(HIBF_GPU_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIBF_TRY_HIP(call, what)                                                                                       \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t const e_ = (call);                                                                                  \
        if (e_ != hipSuccess)                                                                                          \
            return hip_fail(e_, what);                                                                                 \
    } while (0)

uint64_t countl_zero64(uint64_t x)
{
    return x == 0 ? 64u : static_cast<uint64_t>(__builtin_clzll(x));
}

// src/interleaved_bloom_filter.cpp:32-53: validation order and geometry.
int make_geometry(uint64_t bins, uint64_t bin_size, uint64_t hash_funs, hibf_gpu_ibf_geometry & g)
{
    if (bins == 0)
        return fail(HIBF_GPU_ERR_BIN_COUNT, "The number of bins must be > 0.");
    if (hash_funs == 0 || hash_funs > 5)
        return fail(HIBF_GPU_ERR_HASH_COUNT, "The number of hash functions must be > 0 and <= 5.");
    if (bin_size == 0)
        return fail(HIBF_GPU_ERR_BIN_SIZE, "The size of a bin must be > 0.");
    if (bins >= (1ULL << 32))
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "bin_count must be < 2^32 on the device path");
    g.bins = bins;
    g.bin_size = bin_size;
    g.hash_funs = hash_funs;
    g.hash_shift = countl_zero64(bin_size);
    g.bin_words = (bins + 63u) / 64u;
    g.technical_bins = g.bin_words * 64u;
    unsigned __int128 const bits = static_cast<unsigned __int128>(g.technical_bins) * bin_size;
    if (bits >> 63)
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "bit matrix too large");
    g.n_words = static_cast<uint64_t>(bits / 64u);
    return HIBF_GPU_OK;
}

IbfView view_of(hibf_gpu_ibf_geometry const & g, uint64_t const * words)
{
    return IbfView{words,
                   g.bin_size,
                   static_cast<uint32_t>(g.bins),
                   static_cast<uint32_t>(g.bin_words),
                   static_cast<uint32_t>(g.hash_shift),
                   static_cast<uint32_t>(g.hash_funs)};
}

int device_cus()
{
    static int cus = 0;
    if (cus == 0)
    {
        int dev = 0;
        hipDeviceProp_t prop{};
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        else
            (void)hipGetLastError();
        if (cus <= 0)
            cus = 256;
    }
    return cus;
}

// Grows a device buffer to at least `bytes`.
int ensure(void *& ptr, size_t & have, size_t bytes, char const * what)
{
    if (have >= bytes && ptr)
        return HIBF_GPU_OK;
    if (ptr)
    {
        HIBF_TRY_HIP(hipFree(ptr), what);
        ptr = nullptr;
        have = 0;
    }
    size_t const want = std::max<size_t>(bytes, 256);
    HIBF_TRY_HIP(hibf_amd::device_malloc(&ptr, want), what);
    have = want;
    return HIBF_GPU_OK;
}

} // namespace

// Workspace of hibf_gpu_hibf_membership_launch_noalloc, the entry point for
// callers that cannot allocate device memory.  One copy per GPU, held from
// when the library's code object is loaded on that GPU until the process
// ends; every call on a GPU shares it, so calls that overlap on one GPU race.
// Layout: the descent counters (256 bytes), the wave kernel's overflow query
// list (8 bytes per query, kNoAllocMaxQueries queries), then the block
// kernel's per-block hit lists.  In host code `scratch` names a host-side
// stand-in: kernels get the device address from hipGetSymbolAddress.
namespace
{
constexpr uint64_t kNoAllocScratchBytes = 64ULL << 20;
constexpr uint64_t kNoAllocCounterBytes = 256;
constexpr uint64_t kNoAllocMaxQueries = 65536;
constexpr uint64_t kNoAllocHitListOffset = kNoAllocCounterBytes + kNoAllocMaxQueries * sizeof(uint64_t);
static_assert(kNoAllocHitListOffset % 256 == 0, "hit lists start 256-byte aligned");
constexpr int kNoAllocMaxDevices = 64;
} // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace
{
std::mutex g_noalloc_mutex;
unsigned char * g_noalloc_scratch[kNoAllocMaxDevices] = {};

// Device address of this GPU's `scratch`, looked up once per device with the
// device current (the first no-allocation call on it).
int noalloc_scratch(unsigned char ** out)
{
    int dev = 0;
    HIBF_TRY_HIP(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kNoAllocMaxDevices)
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "no-allocation workspace: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_noalloc_mutex);
    if (!g_noalloc_scratch[dev])
    {
        void * p = nullptr;
        HIBF_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p)
            return fail(HIBF_GPU_ERR_DEVICE, "hipGetSymbolAddress(scratch) returned null");
        g_noalloc_scratch[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_noalloc_scratch[dev];
    return HIBF_GPU_OK;
}
} // namespace

namespace hibf_amd
{
// Records the detail message of a failing call (hibf_gpu_last_error); used by
// the host files of the library that are not part of this one.
int set_error(int status, std::string msg)
{
    return fail(status, std::move(msg));
}
} // namespace hibf_amd

struct hibf_gpu_ibf_s
{
    hibf_gpu_ibf_geometry g{};
    uint64_t * d_words{nullptr};
    std::vector<uint64_t> occupancy; // host copy, technical_bins entries (interleaved_bloom_filter.hpp:451-458)
    bool track_occupancy{false};
    bool owns_words{true}; // false: a view of one IBF in an HIBF's arena
};

struct hibf_gpu_agent_s
{
    uint64_t max_queries{65536};
    void * scratch{nullptr}; // HIBF membership: per-block hit lists
    size_t scratch_bytes{0};
    void * ids_tmp{nullptr};
    size_t ids_tmp_bytes{0};
    void * hit_start{nullptr};
    size_t hit_start_bytes{0};
    void * hit_count{nullptr};
    size_t hit_count_bytes{0};
    void * incl{nullptr};
    size_t incl_bytes{0};
    void * cub_temp{nullptr};
    size_t cub_temp_bytes{0};
    void * counts{nullptr}; // IBF membership: uint16 counts of one chunk
    size_t counts_bytes{0};
    void * mask{nullptr}; // IBF membership: count >= threshold bits of one chunk
    size_t mask_bytes{0};
    void * overflow{nullptr}; // HIBF membership: queries answered by the block kernel
    size_t overflow_bytes{0};
    unsigned long long * d_total{nullptr}; // 8 counters, see HibfLaunch::counters
    unsigned long long * h_total{nullptr}; // pinned, 8 entries
};

struct hibf_gpu_hibf_s
{
    std::vector<hibf_gpu_ibf_geometry> geoms;
    std::vector<uint64_t> word_offsets;
    std::vector<uint64_t> tb_offsets;
    std::vector<uint64_t> next_ibf;
    std::vector<uint64_t> user_bin;
    uint64_t n_words{0};
    uint64_t n_user_bins{0};
    uint32_t max_bins{0};
    uint32_t depth{0};
    uint32_t uniform_hash_funs{0}; // hash-function count shared by every IBF, 0 if they differ
    // Run-end TBs of the IBFs reachable from IBF 0: the most ids one query can
    // produce (each IBF is visited at most once; a user bin whose run is cut by
    // a deleted TB, or that has two runs, is reported once per run).
    uint64_t max_hits{0};
    // Some user-bin id occurs in two IBFs.  The reference never builds such
    // tables; bulk_count refuses them (its device scan stores an IBF's counts
    // before descending into its merged bins, the reference in bin order).
    bool user_bin_in_two_ibfs{false};
    // Words of device memory per descent block for counters and event lists
    // ((1 + depth) x max_bins) when they do not fit the LDS budget; 0 when
    // they live in LDS.
    uint64_t gstate_words{0};
    // Phase times of the build that made this HIBF (seconds): sketches and
    // their estimates, layout, allocation (arena and tables), device fill.
    double build_times[4]{};
    double layout_times[2]{}; // union estimation and rearrangement (parts of build_times[1])
    std::string layout_text;  // the layout the build used (reference text format); empty if not built here
    uint64_t * d_arena{nullptr};
    HibfIbfDesc * d_desc{nullptr};
    uint64_t * d_tb_info{nullptr};
};

extern "C" {

const char * hibf_gpu_status_string(int status)
{
    switch (status)
    {
    case HIBF_GPU_OK: return "ok";
    case HIBF_GPU_ERR_INVALID_ARGUMENT: return "invalid argument";
    case HIBF_GPU_ERR_BIN_COUNT: return "The number of bins must be > 0.";
    case HIBF_GPU_ERR_HASH_COUNT: return "The number of hash functions must be > 0 and <= 5.";
    case HIBF_GPU_ERR_BIN_SIZE: return "The size of a bin must be > 0.";
    case HIBF_GPU_ERR_OUT_OF_MEMORY: return "device out of memory";
    case HIBF_GPU_ERR_DEVICE: return "HIP runtime error";
    case HIBF_GPU_ERR_CAPACITY: return "output buffer too small";
    case HIBF_GPU_ERR_UNSUPPORTED: return "unsupported geometry";
    case HIBF_GPU_ERR_CONFIG: return "invalid config";
    case HIBF_GPU_ERR_BUDGET: return "over the counter-memory budget";
    case HIBF_GPU_ERR_NO_DEVICE: return "no HIP device";
    case HIBF_GPU_ERR_IO: return "file or archive error";
    default: return "unknown status";
    }
}

const char * hibf_gpu_last_error(void)
{
    return g_last_error.c_str();
}

int hibf_gpu_tune_set(char const * key, int64_t value)
{
    if (!key)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null key");
    std::string const k{key};
    if (k == "count.wide_kernel" && (value == 0 || value == 1))
        g_count_tuning.wide_kernel = static_cast<uint32_t>(value);
    else if (k == "count.wide_unroll" && (value == 1 || value == 2 || value == 4 || value == 8))
        g_count_tuning.wide_unroll = static_cast<uint32_t>(value);
    else if (k == "count.wide_waves" && value >= 1 && value <= 4)
        g_count_tuning.wide_waves = static_cast<uint32_t>(value);
    else if (k == "hibf.wave_kernel" && (value == 0 || value == 1))
        g_count_tuning.hibf_wave = static_cast<uint32_t>(value);
    else if (k == "hibf.nt_loads" && (value == 0 || value == 1))
        g_count_tuning.nt_loads = static_cast<uint32_t>(value);
    else if (k == "count.nt_stores" && (value == 0 || value == 1))
        g_count_tuning.nt_stores = static_cast<uint32_t>(value);
    else if (k == "count.wide_nt_loads" && (value == 0 || value == 1))
        g_count_tuning.wide_nt_loads = static_cast<uint32_t>(value);
    else if (k == "hibf.stream_kib" && value >= 0 && value <= 0x7FFFFFFF)
        g_count_tuning.hibf_stream_kib = static_cast<uint32_t>(value);
    else if (k == "contains.unroll" && (value == 1 || value == 2 || value == 4 || value == 8))
        g_count_tuning.contains_unroll = static_cast<uint32_t>(value);
    else if (k == "hibf.block_waves" && (value == 1 || value == 2 || value == 4 || value == 8))
        g_count_tuning.hibf_block_waves = static_cast<uint32_t>(value);
    else if (k == "count.counter_bits" && (value == 8 || value == 16))
        g_count_tuning.counter_bits = static_cast<uint32_t>(value);
    else if (k == "hibf.row_per_lane" && (value == 0 || value == 1))
        g_count_tuning.row_per_lane = static_cast<uint32_t>(value);
    else if (k == "hibf.sliced_counts" && (value == 0 || value == 1))
        g_count_tuning.sliced_counts = static_cast<uint32_t>(value);
    else if (k == "hibf.unroll" && (value == 2 || value == 4))
        g_count_tuning.hibf_unroll = static_cast<uint32_t>(value);
    else if (k == "alloc.max_bytes" && value >= 0)
        g_alloc_max_bytes = static_cast<uint64_t>(value);
    else if (k == "build.split_order" && (value == 0 || value == 1))
        g_count_tuning.split_order = static_cast<uint32_t>(value);
    else
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "unknown tuning key or value out of range: " + k);
    return HIBF_GPU_OK;
}

int hibf_gpu_tune_get(char const * key, int64_t * value)
{
    if (!key || !value)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    std::string const k{key};
    if (k == "count.wide_kernel")
        *value = g_count_tuning.wide_kernel;
    else if (k == "count.wide_unroll")
        *value = g_count_tuning.wide_unroll;
    else if (k == "count.wide_waves")
        *value = g_count_tuning.wide_waves;
    else if (k == "hibf.wave_kernel")
        *value = g_count_tuning.hibf_wave;
    else if (k == "hibf.nt_loads")
        *value = g_count_tuning.nt_loads;
    else if (k == "count.nt_stores")
        *value = g_count_tuning.nt_stores;
    else if (k == "count.wide_nt_loads")
        *value = g_count_tuning.wide_nt_loads;
    else if (k == "hibf.stream_kib")
        *value = g_count_tuning.hibf_stream_kib;
    else if (k == "contains.unroll")
        *value = g_count_tuning.contains_unroll;
    else if (k == "hibf.block_waves")
        *value = g_count_tuning.hibf_block_waves;
    else if (k == "count.counter_bits")
        *value = g_count_tuning.counter_bits;
    else if (k == "hibf.row_per_lane")
        *value = g_count_tuning.row_per_lane;
    else if (k == "hibf.sliced_counts")
        *value = g_count_tuning.sliced_counts;
    else if (k == "hibf.unroll")
        *value = g_count_tuning.hibf_unroll;
    else if (k == "alloc.max_bytes")
        *value = static_cast<int64_t>(g_alloc_max_bytes);
    else if (k == "build.split_order")
        *value = g_count_tuning.split_order;
    else
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "unknown tuning key: " + k);
    return HIBF_GPU_OK;
}

int hibf_gpu_malloc(void ** d_ptr, uint64_t bytes)
{
    if (!d_ptr)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    *d_ptr = nullptr;
    HIBF_TRY_HIP(hibf_amd::device_malloc(d_ptr, std::max<uint64_t>(bytes, 1)), "hibf_gpu_malloc");
    return HIBF_GPU_OK;
}

int hibf_gpu_free(void * d_ptr)
{
    if (d_ptr)
        HIBF_TRY_HIP(hipFree(d_ptr), "hibf_gpu_free");
    return HIBF_GPU_OK;
}

int hibf_gpu_memcpy(void * dst, void const * src, uint64_t bytes, int kind)
{
    if ((!dst || !src) && bytes)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (bytes == 0)
        return HIBF_GPU_OK;
    hipMemcpyKind const k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    HIBF_TRY_HIP(hipMemcpy(dst, src, bytes, k), "hibf_gpu_memcpy");
    return HIBF_GPU_OK;
}

int hibf_gpu_synchronize(void)
{
    HIBF_TRY_HIP(hipDeviceSynchronize(), "hibf_gpu_synchronize");
    return HIBF_GPU_OK;
}

/* ---------------------------------------------------------------- IBF --- */

int hibf_gpu_ibf_create(uint64_t bins, uint64_t bin_size, uint32_t hash_funs, hibf_gpu_ibf * out)
{
    if (!out)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    hibf_gpu_ibf_geometry g{};
    if (int const rc = make_geometry(bins, bin_size, hash_funs, g))
        return rc;
    auto * h = new (std::nothrow) hibf_gpu_ibf_s{};
    if (!h)
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    h->g = g;
    hipError_t e = hibf_amd::device_malloc(reinterpret_cast<void **>(&h->d_words), g.n_words * sizeof(uint64_t));
    if (e == hipSuccess)
        e = hipMemset(h->d_words, 0, g.n_words * sizeof(uint64_t));
    if (e != hipSuccess)
    {
        if (h->d_words)
            (void)hipFree(h->d_words);
        delete h;
        return hip_fail(e, "hibf_gpu_ibf_create");
    }
    try
    {
        h->occupancy.assign(g.technical_bins, 0u);
    }
    catch (...)
    {
        (void)hipFree(h->d_words);
        delete h;
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    *out = h;
    return HIBF_GPU_OK;
}

int hibf_gpu_ibf_occupancy(hibf_gpu_ibf ibf, uint64_t * h_occupancy, int * track_occupancy)
{
    if (!ibf)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null handle");
    if (h_occupancy)
        std::copy(ibf->occupancy.begin(), ibf->occupancy.end(), h_occupancy);
    if (track_occupancy)
        *track_occupancy = ibf->track_occupancy ? 1 : 0;
    return HIBF_GPU_OK;
}

int hibf_gpu_ibf_set_occupancy(hibf_gpu_ibf ibf, uint64_t const * h_occupancy, int track_occupancy)
{
    if (!ibf)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null handle");
    if (h_occupancy)
        std::copy(h_occupancy, h_occupancy + ibf->occupancy.size(), ibf->occupancy.begin());
    else
        std::fill(ibf->occupancy.begin(), ibf->occupancy.end(), 0u);
    ibf->track_occupancy = track_occupancy != 0;
    return HIBF_GPU_OK;
}

int hibf_gpu_ibf_destroy(hibf_gpu_ibf ibf)
{
    if (!ibf)
        return HIBF_GPU_OK;
    hipError_t const e = (ibf->d_words && ibf->owns_words) ? hipFree(ibf->d_words) : hipSuccess;
    delete ibf;
    return e == hipSuccess ? HIBF_GPU_OK : hip_fail(e, "hibf_gpu_ibf_destroy");
}

int hibf_gpu_ibf_get_geometry(hibf_gpu_ibf ibf, hibf_gpu_ibf_geometry * out)
{
    if (!ibf || !out)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    *out = ibf->g;
    return HIBF_GPU_OK;
}

int hibf_gpu_ibf_device_words(hibf_gpu_ibf ibf, uint64_t ** d_words)
{
    if (!ibf || !d_words)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    *d_words = ibf->d_words;
    return HIBF_GPU_OK;
}

int hibf_gpu_ibf_upload(hibf_gpu_ibf ibf, uint64_t const * h_words, uint64_t n_words)
{
    if (!ibf || (!h_words && n_words))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (n_words != ibf->g.n_words)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "word count does not match the geometry");
    HIBF_TRY_HIP(hipMemcpy(ibf->d_words, h_words, n_words * sizeof(uint64_t), hipMemcpyHostToDevice),
                 "hibf_gpu_ibf_upload");
    return HIBF_GPU_OK;
}

int hibf_gpu_ibf_download(hibf_gpu_ibf ibf, uint64_t * h_words, uint64_t n_words)
{
    if (!ibf || (!h_words && n_words))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (n_words != ibf->g.n_words)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "word count does not match the geometry");
    HIBF_TRY_HIP(hipDeviceSynchronize(), "hibf_gpu_ibf_download");
    HIBF_TRY_HIP(hipMemcpy(h_words, ibf->d_words, n_words * sizeof(uint64_t), hipMemcpyDeviceToHost),
                 "hibf_gpu_ibf_download");
    return HIBF_GPU_OK;
}

namespace
{
// Emplace with track_occupancy: occupancy[bin] grows by one for each pair (in
// the given order) that sets a bit unset before it, as n calls of the
// reference's emplace do (src/interleaved_bloom_filter.cpp:119-140).  The
// pairs' positions are sorted stably; a pair counts iff it is the first to
// touch one of its positions and that bit was unset before the call.
// Synchronous: the counts are added to the host copy of the occupancy.
int emplace_tracked(hibf_gpu_ibf ibf, uint64_t const * d_values, uint64_t const * d_bins, uint64_t bin, uint64_t n,
                    hipStream_t s)
{
    hibf_gpu_ibf_geometry const & g = ibf->g;
    uint64_t const m = n * g.hash_funs;
    if (n > (1ULL << 30) || m > static_cast<uint64_t>(INT32_MAX))
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "emplace with track_occupancy: at most 2^31 - 1 positions per call");
    uint64_t const invalid_key = g.technical_bins * g.bin_size; // one past the last bit
    int end_bit = 1;
    while (end_bit < 64 && (invalid_key >> end_bit) != 0u)
        ++end_bit;
    size_t sort_bytes = 0;
    HIBF_TRY_HIP(radix_sort_pairs_u64(nullptr, sort_bytes, nullptr, nullptr, nullptr, nullptr, m, end_bit, s),
                 "tracked emplace sort size");
    uint64_t const occ_bytes = g.technical_bins * sizeof(unsigned long long);
    uint64_t const pair_bytes = m * 8u;
    // keys, sorted keys, owners, sorted owners, marks, per-bin counts, sort temp
    uint64_t const off_keys2 = pair_bytes, off_own = 2u * pair_bytes, off_own2 = 3u * pair_bytes;
    uint64_t const off_marks = 4u * pair_bytes;
    uint64_t const off_occ = (off_marks + n + 255u) / 256u * 256u;
    uint64_t const off_tmp = (off_occ + occ_bytes + 255u) / 256u * 256u;
    char * ws = nullptr;
    HIBF_TRY_HIP(hibf_amd::device_malloc(&ws, off_tmp + sort_bytes), "tracked emplace workspace");
    std::unique_ptr<char, void (*)(char *)> guard{ws, [](char * p) { (void)hipFree(p); }};
    auto * keys = reinterpret_cast<uint64_t *>(ws);
    auto * keys2 = reinterpret_cast<uint64_t *>(ws + off_keys2);
    auto * owners = reinterpret_cast<uint64_t *>(ws + off_own);
    auto * owners2 = reinterpret_cast<uint64_t *>(ws + off_own2);
    auto * marks = reinterpret_cast<uint8_t *>(ws + off_marks);
    auto * occ = reinterpret_cast<unsigned long long *>(ws + off_occ);
    HIBF_TRY_HIP(hipMemsetAsync(marks, 0, n, s), "tracked emplace marks");
    HIBF_TRY_HIP(hipMemsetAsync(occ, 0, occ_bytes, s), "tracked emplace counts");
    HIBF_TRY_HIP(launch_emplace_positions(g.bin_size, static_cast<uint32_t>(g.bin_words),
                                          static_cast<uint32_t>(g.hash_shift), static_cast<uint32_t>(g.hash_funs),
                                          d_values, d_bins, bin, n, g.bins, invalid_key, keys, owners, s),
                 "tracked emplace positions");
    HIBF_TRY_HIP(radix_sort_pairs_u64(ws + off_tmp, sort_bytes, keys, keys2, owners, owners2, m, end_bit, s),
                 "tracked emplace sort");
    HIBF_TRY_HIP(launch_mark_first_touches(ibf->d_words, keys2, owners2, m, invalid_key, marks, s),
                 "tracked emplace first touches");
    HIBF_TRY_HIP(launch_emplace(ibf->d_words, g.bin_size, static_cast<uint32_t>(g.bin_words),
                                static_cast<uint32_t>(g.hash_shift), static_cast<uint32_t>(g.hash_funs), d_values,
                                d_bins, bin, n, g.bins, s),
                 "emplace launch");
    HIBF_TRY_HIP(launch_count_marked(marks, d_bins, bin, n, occ, s), "tracked emplace counts");
    std::vector<uint64_t> add(g.technical_bins);
    HIBF_TRY_HIP(hipMemcpyAsync(add.data(), occ, occ_bytes, hipMemcpyDeviceToHost, s), "tracked emplace copy");
    HIBF_TRY_HIP(hipStreamSynchronize(s), "tracked emplace sync");
    for (uint64_t b = 0; b < g.technical_bins && b < ibf->occupancy.size(); ++b)
        ibf->occupancy[b] += add[b];
    return HIBF_GPU_OK;
}
} // namespace

int hibf_gpu_ibf_emplace(hibf_gpu_ibf ibf,
                         uint64_t const * d_values,
                         uint64_t const * d_bins,
                         uint64_t bin,
                         uint64_t n,
                         void * stream)
{
    if (!ibf || (!d_values && n))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (!d_bins && bin >= ibf->g.bins)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "bin index >= bin_count");
    if (ibf->track_occupancy && n)
        return emplace_tracked(ibf, d_values, d_bins, bin, n, static_cast<hipStream_t>(stream));
    HIBF_TRY_HIP(launch_emplace(ibf->d_words, ibf->g.bin_size, static_cast<uint32_t>(ibf->g.bin_words),
                                static_cast<uint32_t>(ibf->g.hash_shift), static_cast<uint32_t>(ibf->g.hash_funs),
                                d_values, d_bins, bin, n, ibf->g.bins, static_cast<hipStream_t>(stream)),
                 "emplace launch");
    return HIBF_GPU_OK;
}

int hibf_gpu_ibf_emplace_synthetic(hibf_gpu_ibf ibf,
                                   uint64_t seed,
                                   uint64_t first_bin,
                                   uint64_t n_bins,
                                   uint64_t values_per_bin,
                                   void * stream)
{
    if (!ibf)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (first_bin > ibf->g.bins || n_bins > ibf->g.bins - first_bin)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "bin range exceeds bin_count");
    if (values_per_bin >= (1ULL << 36))
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "values_per_bin must be < 2^36");
    HIBF_TRY_HIP(launch_emplace_synth(ibf->d_words, ibf->g.bin_size, static_cast<uint32_t>(ibf->g.bin_words),
                                      static_cast<uint32_t>(ibf->g.hash_shift),
                                      static_cast<uint32_t>(ibf->g.hash_funs), seed, first_bin, n_bins,
                                      values_per_bin, static_cast<hipStream_t>(stream)),
                 "emplace_synthetic launch");
    return HIBF_GPU_OK;
}

uint64_t hibf_gpu_synth_value(uint64_t seed, uint64_t bin, uint64_t i)
{
    return synth_value(seed, bin, i);
}

int hibf_gpu_ibf_clear(hibf_gpu_ibf ibf, uint64_t const * h_bins, uint64_t n, void * stream)
{
    if (!ibf || (!h_bins && n))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0)
        return HIBF_GPU_OK;
    for (uint64_t i = 0; i < n; ++i)
        if (h_bins[i] >= ibf->g.bins)
            return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "bin index >= bin_count");
    if (n >= (1ULL << 31))
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "too many bins");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    uint64_t * d_bins = nullptr;
    HIBF_TRY_HIP(hibf_amd::device_malloc(reinterpret_cast<void **>(&d_bins), n * sizeof(uint64_t)), "clear alloc");
    hipError_t e = hipMemcpyAsync(d_bins, h_bins, n * sizeof(uint64_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = launch_clear(ibf->d_words, ibf->g.bin_size, static_cast<uint32_t>(ibf->g.bin_words), d_bins,
                         static_cast<uint32_t>(n), s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    (void)hipFree(d_bins);
    if (e != hipSuccess)
        return hip_fail(e, "hibf_gpu_ibf_clear");
    for (uint64_t i = 0; i < n; ++i) // src/interleaved_bloom_filter.cpp:145
        ibf->occupancy[h_bins[i]] = 0u;
    return HIBF_GPU_OK;
}

int hibf_gpu_ibf_try_increase_bin_number_to(hibf_gpu_ibf ibf, uint64_t new_bins, int * increased)
{
    if (!ibf || !increased)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (!ibf->owns_words)
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "the IBFs of an HIBF keep the bin counts of its routing tables");
    uint64_t const new_bw = (new_bins + 63u) / 64u;
    *increased = 0;
    if (new_bins < ibf->g.bins || new_bw > ibf->g.bin_words) // src/interleaved_bloom_filter.cpp:150-160
        return HIBF_GPU_OK;
    if (new_bins >= (1ULL << 32))
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "bin_count must be < 2^32 on the device path");
    ibf->g.bins = new_bins;
    *increased = 1;
    return HIBF_GPU_OK;
}

int hibf_gpu_ibf_increase_bin_number_to(hibf_gpu_ibf ibf, uint64_t new_bins)
{
    if (!ibf)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null handle");
    if (new_bins < ibf->g.bins) // src/interleaved_bloom_filter.cpp:165-166
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "The number of new bins must be >= the current number of bins.");
    int same_words = 0;
    if (int rc = hibf_gpu_ibf_try_increase_bin_number_to(ibf, new_bins, &same_words))
        return rc;
    if (same_words)
        return HIBF_GPU_OK;
    hibf_gpu_ibf_geometry g{};
    if (int rc = make_geometry(new_bins, ibf->g.bin_size, ibf->g.hash_funs, g))
        return rc;
    std::vector<uint64_t> occ;
    try
    {
        occ.assign(g.technical_bins, 0u);
    }
    catch (...)
    {
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    std::copy(ibf->occupancy.begin(), ibf->occupancy.end(), occ.begin());
    uint64_t * d_new = nullptr;
    HIBF_TRY_HIP(hibf_amd::device_malloc(reinterpret_cast<void **>(&d_new), g.n_words * sizeof(uint64_t)), "increase_bin_number_to");
    hipError_t e = launch_restride(ibf->d_words, ibf->g.bin_words, d_new, g.bin_words, g.bin_size, nullptr);
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    if (e != hipSuccess)
    {
        (void)hipFree(d_new);
        return hip_fail(e, "increase_bin_number_to");
    }
    (void)hipFree(ibf->d_words);
    ibf->d_words = d_new;
    ibf->g = g;
    ibf->occupancy.swap(occ);
    return HIBF_GPU_OK;
}

int hibf_gpu_ibf_bulk_contains(hibf_gpu_ibf ibf, uint64_t const * d_values, uint64_t n, uint64_t * d_out, void * stream)
{
    if (!ibf || ((!d_values || !d_out) && n))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    HIBF_TRY_HIP(launch_contains(view_of(ibf->g, ibf->d_words), d_values, n, d_out, static_cast<hipStream_t>(stream)),
                 "bulk_contains launch");
    return HIBF_GPU_OK;
}

int hibf_gpu_ibf_bulk_count(hibf_gpu_ibf ibf,
                            uint64_t const * d_values,
                            uint64_t const * d_offsets,
                            uint64_t nq,
                            uint32_t counter_bytes,
                            void * d_counts,
                            void * stream)
{
    if (!ibf || ((!d_offsets || !d_counts) && nq))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (counter_bytes != 1 && counter_bytes != 2 && counter_bytes != 4 && counter_bytes != 8)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "counter_bytes must be 1, 2, 4 or 8");
    HIBF_TRY_HIP(launch_count(view_of(ibf->g, ibf->d_words), d_values, d_offsets, nq, counter_bytes, d_counts,
                              1u << 24, static_cast<hipStream_t>(stream)),
                 "bulk_count launch");
    return HIBF_GPU_OK;
}

/* -------------------------------------------------------------- agent --- */

int hibf_gpu_agent_create(uint64_t max_queries_per_launch, hibf_gpu_agent * out)
{
    if (!out)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    auto * a = new (std::nothrow) hibf_gpu_agent_s{};
    if (!a)
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    a->max_queries = max_queries_per_launch ? max_queries_per_launch : 65536u;
    hipError_t e = hibf_amd::device_malloc(reinterpret_cast<void **>(&a->d_total), 8 * sizeof(unsigned long long));
    if (e == hipSuccess)
        e = hipMemset(a->d_total, 0, 8 * sizeof(unsigned long long));
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void **>(&a->h_total), 8 * sizeof(unsigned long long), hipHostMallocDefault);
    if (e != hipSuccess)
    {
        if (a->d_total)
            (void)hipFree(a->d_total);
        delete a;
        return hip_fail(e, "hibf_gpu_agent_create");
    }
    *out = a;
    return HIBF_GPU_OK;
}

int hibf_gpu_agent_destroy(hibf_gpu_agent a)
{
    if (!a)
        return HIBF_GPU_OK;
    for (void * p : {a->scratch, a->ids_tmp, a->hit_start, a->hit_count, a->incl, a->cub_temp, a->counts, a->mask,
                     a->overflow,
                     static_cast<void *>(a->d_total)})
        if (p)
            (void)hipFree(p);
    if (a->h_total)
        (void)hipHostFree(a->h_total);
    delete a;
    return HIBF_GPU_OK;
}

int hibf_gpu_agent_stats(hibf_gpu_agent a, uint64_t * visits, uint64_t * lookups, uint64_t * row_bytes, int reset)
{
    if (!a)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null agent");
    HIBF_TRY_HIP(hipDeviceSynchronize(), "stats sync");
    HIBF_TRY_HIP(hipMemcpy(a->h_total, a->d_total, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost), "stats copy");
    if (visits)
        *visits = a->h_total[2];
    if (row_bytes)
        *row_bytes = a->h_total[3];
    if (lookups)
        *lookups = a->h_total[4];
    if (reset)
        HIBF_TRY_HIP(hipMemset(a->d_total + 2, 0, 4 * sizeof(unsigned long long)), "stats reset");
    return HIBF_GPU_OK;
}

int hibf_gpu_agent_status(hibf_gpu_agent a, uint64_t * last_total, uint64_t * dropped)
{
    if (!a)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null agent");
    HIBF_TRY_HIP(hipDeviceSynchronize(), "status sync");
    HIBF_TRY_HIP(hipMemcpy(a->h_total, a->d_total, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost), "status copy");
    if (last_total)
        *last_total = a->h_total[0];
    if (dropped)
        *dropped = a->h_total[5];
    return HIBF_GPU_OK;
}

namespace
{

// d_hit_offsets[q0 .. q0+n] from hit_count[0..n) (write_first: also d_hit_offsets[q0] = base).
// d_offsets[1 + i] = *d_base (0 when d_base is null) + inclusive sum of
// d_count[0 .. i]; d_offsets[0] is written (0) only when d_base is null.
int build_offsets(hibf_gpu_agent a,
                  uint64_t const * d_count,
                  uint64_t n,
                  uint64_t const * d_base,
                  uint64_t * d_offsets,
                  hipStream_t s)
{
    if (int rc = ensure(a->incl, a->incl_bytes, n * sizeof(uint64_t), "offsets alloc"))
        return rc;
    size_t need = 0;
    HIBF_TRY_HIP(inclusive_sum_u64(nullptr, need, d_count, static_cast<uint64_t *>(a->incl), n, s), "scan size");
    if (int rc = ensure(a->cub_temp, a->cub_temp_bytes, need, "scan alloc"))
        return rc;
    size_t have = a->cub_temp_bytes;
    HIBF_TRY_HIP(inclusive_sum_u64(a->cub_temp, have, d_count, static_cast<uint64_t *>(a->incl), n, s), "scan");
    HIBF_TRY_HIP(launch_finalize_offsets(static_cast<uint64_t *>(a->incl), n, 0, d_base, d_base == nullptr, d_offsets,
                                         s),
                 "offsets");
    return HIBF_GPU_OK;
}

} // namespace

int hibf_gpu_ibf_membership_for(hibf_gpu_ibf ibf,
                                hibf_gpu_agent a,
                                uint64_t const * d_values,
                                uint64_t const * d_offsets,
                                uint64_t nq,
                                uint16_t const * d_thresholds,
                                uint64_t * d_ids,
                                uint64_t ids_cap,
                                uint64_t * d_hit_offsets,
                                uint64_t * h_total,
                                void * stream)
{
    if (!ibf || !a || !h_total || ((!d_offsets || !d_thresholds || !d_hit_offsets) && nq))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    *h_total = 0;
    if (nq == 0)
    {
        return HIBF_GPU_OK;
    }
    // Chunk the queries so that the uint16 counts of one chunk stay within 256 MiB.  The chunks are queued
    // back to back: each one's offsets continue from the previous chunk's last offset in device memory, so
    // the host waits once, for the total.
    uint64_t const bins = ibf->g.bins;
    uint64_t const mask_words = ibf->g.bin_words;
    uint64_t chunk = std::max<uint64_t>(1u, (256ULL << 20) / (bins * 2u));
    chunk = std::min<uint64_t>(chunk, std::min<uint64_t>(nq, a->max_queries));
    if (int rc = ensure(a->counts, a->counts_bytes, chunk * bins * 2u, "membership counts"))
        return rc;
    if (int rc = ensure(a->hit_count, a->hit_count_bytes, chunk * sizeof(uint64_t), "membership counts"))
        return rc;
    if (int rc = ensure(a->mask, a->mask_bytes, chunk * mask_words * sizeof(uint64_t), "membership masks"))
        return rc;
    IbfView const g = view_of(ibf->g, ibf->d_words);
    auto * counts = static_cast<uint16_t *>(a->counts);
    auto * hc = static_cast<uint64_t *>(a->hit_count);
    auto * mask = static_cast<uint64_t *>(a->mask);
    for (uint64_t q0 = 0; q0 < nq; q0 += chunk)
    {
        uint64_t const n = std::min(chunk, nq - q0);
        HIBF_TRY_HIP(launch_count(g, d_values, d_offsets + q0, n, 2u, counts, 1u << 24, s), "count");
        HIBF_TRY_HIP(launch_threshold_mask(counts, bins, n, d_thresholds + q0, hc, mask, mask_words, s),
                     "threshold mask");
        if (int rc = build_offsets(a, hc, n, q0 == 0 ? nullptr : d_hit_offsets + q0, d_hit_offsets + q0, s))
            return rc;
        HIBF_TRY_HIP(launch_mask_compact(mask, mask_words, n, d_hit_offsets + q0, d_ids, d_ids ? ids_cap : 0, s),
                     "mask compaction");
    }
    HIBF_TRY_HIP(hipMemcpyAsync(a->h_total, d_hit_offsets + nq, sizeof(uint64_t), hipMemcpyDeviceToHost, s),
                 "total copy");
    HIBF_TRY_HIP(hipStreamSynchronize(s), "membership sync");
    uint64_t const total = *a->h_total;
    *h_total = total;
    if (total > ids_cap || (!d_ids && total))
        return fail(HIBF_GPU_ERR_CAPACITY, "ids_cap too small; *h_total holds the needed size");
    return HIBF_GPU_OK;
}

/* --------------------------------------------------------------- HIBF --- */

int hibf_gpu_hibf_create(uint64_t n_ibfs,
                         hibf_gpu_ibf_geometry const * geoms,
                         uint64_t const * h_arena,
                         uint64_t const * h_word_offsets,
                         uint64_t const * h_tb_offsets,
                         uint64_t const * h_next_ibf,
                         uint64_t const * h_user_bin,
                         uint64_t number_of_user_bins,
                         hibf_gpu_hibf * out)
{
    if (!out || !geoms || !h_word_offsets || !h_tb_offsets || !h_next_ibf || !h_user_bin || n_ibfs == 0)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument or no IBF");
    *out = nullptr;
    if (n_ibfs >= (1ULL << 31) || number_of_user_bins >= (1ULL << 31))
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "too many IBFs or user bins for 31-bit device ids");
    auto * h = new (std::nothrow) hibf_gpu_hibf_s{};
    if (!h)
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    std::unique_ptr<hibf_gpu_hibf_s> guard{h};
    h->n_user_bins = number_of_user_bins;
    h->geoms.resize(n_ibfs);
    h->word_offsets.assign(h_word_offsets, h_word_offsets + n_ibfs);
    h->tb_offsets.assign(h_tb_offsets, h_tb_offsets + n_ibfs + 1);
    h->next_ibf.assign(h_next_ibf, h_next_ibf + h_tb_offsets[n_ibfs]);
    h->user_bin.assign(h_user_bin, h_user_bin + h_tb_offsets[n_ibfs]);
    std::vector<HibfIbfDesc> desc(n_ibfs);
    uint64_t const n_tbs = h_tb_offsets[n_ibfs];
    std::vector<uint64_t> info(n_tbs);
    uint64_t end_words = 0;
    for (uint64_t i = 0; i < n_ibfs; ++i)
    {
        hibf_gpu_ibf_geometry g{};
        if (int rc = make_geometry(geoms[i].bins, geoms[i].bin_size, geoms[i].hash_funs, g))
            return rc;
        h->geoms[i] = g;
        if (h_tb_offsets[i + 1] - h_tb_offsets[i] != g.bins)
            return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "table size of an IBF differs from its bin_count");
        if (h_word_offsets[i] % 4u)
            return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "IBF word offsets must be multiples of 4 (32-byte aligned)");
        end_words = std::max(end_words, h_word_offsets[i] + g.n_words);
        uint32_t lpv_log, cw, tile;
        count_shape(static_cast<uint32_t>(g.bin_words), lpv_log, cw, tile);
        desc[i] = HibfIbfDesc{h_word_offsets[i],
                              g.bin_size,
                              h_tb_offsets[i],
                              static_cast<uint32_t>(g.bins),
                              static_cast<uint32_t>(g.bin_words),
                              static_cast<uint32_t>(g.hash_shift),
                              static_cast<uint32_t>(g.hash_funs),
                              cw == 2u ? 7u : lpv_log,
                              0u};
        h->max_bins = std::max<uint32_t>(h->max_bins, static_cast<uint32_t>(g.bins));
        if (i == 0)
            h->uniform_hash_funs = static_cast<uint32_t>(g.hash_funs);
        else if (h->uniform_hash_funs != g.hash_funs)
            h->uniform_hash_funs = 0u;
        uint64_t const * ub = h_user_bin + h_tb_offsets[i];
        uint64_t const * nx = h_next_ibf + h_tb_offsets[i];
        for (uint64_t b = 0; b < g.bins; ++b)
        {
            uint64_t kind, payload;
            if (ub[b] == HIBF_GPU_BIN_MERGED)
            {
                if (nx[b] >= n_ibfs || nx[b] == 0)
                    return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "merged bin points to an invalid IBF");
                kind = kTbMerged;
                payload = nx[b];
            }
            else if (ub[b] == HIBF_GPU_BIN_DELETED)
            {
                kind = kTbDeleted;
                payload = 0;
            }
            else
            {
                if (ub[b] >= number_of_user_bins)
                    return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "user bin id >= number_of_user_bins");
                bool const end = (b + 1 == g.bins) || ub[b + 1] != ub[b];
                kind = end ? kTbRunEnd : kTbRegular;
                payload = ub[b];
            }
            info[h_tb_offsets[i] + b] = (kind << 32) | payload;
        }
    }
    // Depth of the tree below IBF 0 (iterative, with a cycle guard).
    {
        std::vector<uint32_t> level(n_ibfs, 0u);
        std::vector<uint64_t> stack{0};
        level[0] = 1;
        uint32_t depth = 1;
        uint64_t visited = 0;
        while (!stack.empty())
        {
            uint64_t const i = stack.back();
            stack.pop_back();
            if (++visited > n_ibfs)
                return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "merged bins do not form a tree");
            for (uint64_t b = h_tb_offsets[i]; b < h_tb_offsets[i + 1]; ++b)
                if ((info[b] >> 32) == kTbRunEnd)
                    ++h->max_hits;
                else if ((info[b] >> 32) == kTbMerged)
                {
                    uint64_t const c = info[b] & 0xFFFFFFFFu;
                    if (level[c])
                        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "merged bins do not form a tree");
                    level[c] = level[i] + 1;
                    depth = std::max(depth, level[c]);
                    stack.push_back(c);
                }
        }
        if (depth > kHibfMaxDepth)
            return fail(HIBF_GPU_ERR_UNSUPPORTED, "HIBF deeper than 16 levels");
        h->depth = depth;
    }
    {
        std::vector<std::pair<uint64_t, uint64_t>> owner; // (user bin, IBF) of every user-bin TB
        owner.reserve(n_tbs);
        for (uint64_t i = 0; i < n_ibfs; ++i)
            for (uint64_t b = h_tb_offsets[i]; b < h_tb_offsets[i + 1]; ++b)
                if ((info[b] >> 32) == kTbRegular || (info[b] >> 32) == kTbRunEnd)
                    owner.emplace_back(info[b] & 0xFFFFFFFFu, i);
        std::sort(owner.begin(), owner.end());
        for (size_t k = 1; k < owner.size(); ++k)
            if (owner[k].first == owner[k - 1].first && owner[k].second != owner[k - 1].second)
                h->user_bin_in_two_ibfs = true;
    }
    if (static_cast<uint64_t>(h->max_bins) * (1u + h->depth) * sizeof(uint32_t) > kHibfLdsStateBytes)
        h->gstate_words = static_cast<uint64_t>(h->max_bins) * (1u + h->depth);
    h->n_words = end_words;
    hipError_t e = hibf_amd::device_malloc(reinterpret_cast<void **>(&h->d_arena), std::max<uint64_t>(end_words, 1) * 8u);
    if (e == hipSuccess)
        e = h_arena ? hipMemcpy(h->d_arena, h_arena, end_words * 8u, hipMemcpyHostToDevice)
                    : hipMemset(h->d_arena, 0, end_words * 8u);
    if (e == hipSuccess)
        e = hibf_amd::device_malloc(reinterpret_cast<void **>(&h->d_desc), n_ibfs * sizeof(HibfIbfDesc));
    if (e == hipSuccess)
        e = hipMemcpy(h->d_desc, desc.data(), n_ibfs * sizeof(HibfIbfDesc), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hibf_amd::device_malloc(reinterpret_cast<void **>(&h->d_tb_info), std::max<uint64_t>(n_tbs, 1) * 8u);
    if (e == hipSuccess)
        e = hipMemcpy(h->d_tb_info, info.data(), n_tbs * 8u, hipMemcpyHostToDevice);
    if (e != hipSuccess)
    {
        for (void * p : {static_cast<void *>(h->d_arena), static_cast<void *>(h->d_desc),
                         static_cast<void *>(h->d_tb_info)})
            if (p)
                (void)hipFree(p);
        return hip_fail(e, "hibf_gpu_hibf_create");
    }
    *out = guard.release();
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_info(hibf_gpu_hibf h,
                       uint64_t * n_ibfs,
                       uint64_t * n_tbs,
                       uint64_t * n_words,
                       uint64_t * n_user_bins,
                       uint64_t * depth)
{
    if (!h)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null handle");
    if (n_ibfs)
        *n_ibfs = h->geoms.size();
    if (n_tbs)
        *n_tbs = h->tb_offsets.back();
    if (n_words)
        *n_words = h->n_words;
    if (n_user_bins)
        *n_user_bins = h->n_user_bins;
    if (depth)
        *depth = h->depth;
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_export_tables(hibf_gpu_hibf h,
                                hibf_gpu_ibf_geometry * geoms,
                                uint64_t * word_offsets,
                                uint64_t * tb_offsets,
                                uint64_t * next_ibf,
                                uint64_t * user_bin)
{
    if (!h || !geoms || !word_offsets || !tb_offsets || !next_ibf || !user_bin)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    std::copy(h->geoms.begin(), h->geoms.end(), geoms);
    std::copy(h->word_offsets.begin(), h->word_offsets.end(), word_offsets);
    std::copy(h->tb_offsets.begin(), h->tb_offsets.end(), tb_offsets);
    std::copy(h->next_ibf.begin(), h->next_ibf.end(), next_ibf);
    std::copy(h->user_bin.begin(), h->user_bin.end(), user_bin);
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_destroy(hibf_gpu_hibf h)
{
    if (!h)
        return HIBF_GPU_OK;
    for (void * p : {static_cast<void *>(h->d_arena), static_cast<void *>(h->d_desc),
                     static_cast<void *>(h->d_tb_info)})
        if (p)
            (void)hipFree(p);
    delete h;
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_emplace(hibf_gpu_hibf h,
                          uint64_t ibf_idx,
                          uint64_t const * d_values,
                          uint64_t const * d_bins,
                          uint64_t bin,
                          uint64_t n,
                          void * stream)
{
    if (!h || ibf_idx >= h->geoms.size() || (!d_values && n))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument or IBF index out of range");
    hibf_gpu_ibf_geometry const & g = h->geoms[ibf_idx];
    if (!d_bins && bin >= g.bins)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "bin index >= bin_count");
    HIBF_TRY_HIP(launch_emplace(h->d_arena + h->word_offsets[ibf_idx], g.bin_size,
                                static_cast<uint32_t>(g.bin_words), static_cast<uint32_t>(g.hash_shift),
                                static_cast<uint32_t>(g.hash_funs), d_values, d_bins, bin, n, g.bins,
                                static_cast<hipStream_t>(stream)),
                 "hibf emplace launch");
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_build_times(hibf_gpu_hibf h, double * h_seconds)
{
    if (!h || !h_seconds)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < 4; ++i)
        h_seconds[i] = h->build_times[i];
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_layout_times(hibf_gpu_hibf h, double * h_seconds)
{
    if (!h || !h_seconds)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    h_seconds[0] = h->layout_times[0];
    h_seconds[1] = h->layout_times[1];
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_ibf_view(hibf_gpu_hibf h, uint64_t i, hibf_gpu_ibf * out)
{
    if (!h || !out)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (i >= h->geoms.size())
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "IBF index out of range");
    auto * v = new (std::nothrow) hibf_gpu_ibf_s{};
    if (!v)
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    v->g = h->geoms[i];
    v->d_words = h->d_arena + h->word_offsets[i];
    v->occupancy.assign(v->g.technical_bins, 0u);
    v->owns_words = false;
    *out = v;
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_device_arena(hibf_gpu_hibf h, uint64_t ** d_arena)
{
    if (!h || !d_arena)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    *d_arena = h->d_arena;
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_arena_words(hibf_gpu_hibf h, uint64_t * n_words)
{
    if (!h || !n_words)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    *n_words = h->n_words;
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_download_arena(hibf_gpu_hibf h, uint64_t * h_arena, uint64_t n_words)
{
    if (!h || (!h_arena && n_words))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (n_words != h->n_words)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "word count does not match the arena");
    HIBF_TRY_HIP(hipDeviceSynchronize(), "download sync");
    HIBF_TRY_HIP(hipMemcpy(h_arena, h->d_arena, n_words * 8u, hipMemcpyDeviceToHost), "download arena");
    return HIBF_GPU_OK;
}

namespace
{
HibfLaunch launch_with(hibf_gpu_hibf h, unsigned long long * counters)
{
    uint64_t const stream_bytes =
        g_count_tuning.hibf_stream_kib ? static_cast<uint64_t>(g_count_tuning.hibf_stream_kib) << 10 : ~0ull;
    return HibfLaunch{h->d_desc,      h->d_arena,   h->d_tb_info, h->max_bins, h->depth, h->n_user_bins,
                      h->max_hits,    (h->max_bins + 63u) / 64u, h->uniform_hash_funs, g_count_tuning.nt_stores,
                      0u,             h->gstate_words,           stream_bytes, counters};
}

HibfLaunch launch_of(hibf_gpu_hibf h, hibf_gpu_agent a)
{
    return launch_with(h, a->d_total);
}

// Blocks of `waves` waves for a query-per-block launch: 32 waves per CU (8 per
// SIMD, what the 64-VGPR descent kernels can hold), fewer when the per-block
// scratch would pass 1 GiB, never more than the queries.
unsigned hibf_blocks(uint64_t nq, uint64_t per_block_scratch_bytes, uint32_t waves = 4u)
{
    uint64_t blocks = static_cast<uint64_t>(device_cus()) * (32u / std::max<uint32_t>(waves, 1u));
    if (per_block_scratch_bytes)
        blocks = std::min<uint64_t>(blocks, std::max<uint64_t>(1u, (1ULL << 30) / per_block_scratch_bytes));
    blocks = std::min<uint64_t>(blocks, nq);
    return static_cast<unsigned>(std::max<uint64_t>(blocks, 1u));
}
} // namespace

namespace
{

constexpr uint32_t kMembershipCounters = (1u << 0) | (1u << 1) | (1u << 6) | (1u << 7);

bool wave_kernel_fits(HibfLaunch const & L)
{
    return g_count_tuning.hibf_wave != 0u && hibf_wave_lds_bytes(L, 4) <= 64u * 1024u;
}

// Per-block scratch of the block-per-query kernel: a hit list with room for
// every id one query can produce, then the counters and event lists of HIBFs
// whose state does not fit LDS.
uint64_t block_hit_list_bytes(hibf_gpu_hibf h)
{
    return (std::max<uint64_t>(h->max_hits, 1u) + h->gstate_words) * sizeof(uint32_t);
}

unsigned overflow_blocks(hibf_gpu_hibf h, uint64_t nq)
{
    return std::min<unsigned>(256u, hibf_blocks(nq, block_hit_list_bytes(h), g_count_tuning.hibf_block_waves));
}

// Workspace of the membership path for batches of up to nq queries.
int prepare_membership(hibf_gpu_hibf h, hibf_gpu_agent a, uint64_t nq)
{
    uint64_t const per_block = block_hit_list_bytes(h);
    HibfLaunch const L = launch_of(h, a);
    unsigned const blocks = wave_kernel_fits(L) ? overflow_blocks(h, nq)
                                                : hibf_blocks(nq, per_block, g_count_tuning.hibf_block_waves);
    if (int rc = ensure(a->scratch, a->scratch_bytes, static_cast<size_t>(blocks) * per_block, "membership scratch"))
        return rc;
    if (int rc = ensure(a->overflow, a->overflow_bytes, std::max<uint64_t>(nq, 1) * 8u, "membership overflow"))
        return rc;
    return HIBF_GPU_OK;
}

// Enqueues one membership pass: wave-per-query kernel plus the block-per-query
// pass over the queries whose ids did not fit the wave's buffer (or the block
// kernel alone when the HIBF's IBFs are too wide for per-wave LDS state).
int enqueue_membership(hibf_gpu_hibf h,
                       hibf_gpu_agent a,
                       uint64_t const * d_values,
                       uint64_t const * d_offsets,
                       uint64_t nq,
                       uint16_t const * d_thresholds,
                       uint64_t * d_ids,
                       uint64_t ids_cap,
                       uint64_t * d_hit_start,
                       uint64_t * d_hit_count,
                       hipStream_t s,
                       uint32_t per_query_cap = 0u)
{
    HibfLaunch L = launch_of(h, a);
    L.per_query_cap = per_query_cap;
    uint64_t const per_block = block_hit_list_bytes(h);
    HIBF_TRY_HIP(launch_reset_counters(a->d_total, kMembershipCounters, s), "counter reset");
    if (wave_kernel_fits(L))
    {
        unsigned const ob = overflow_blocks(h, nq);
        if (a->scratch_bytes < static_cast<size_t>(ob) * per_block || a->overflow_bytes < nq * 8u)
            return fail(HIBF_GPU_ERR_INVALID_ARGUMENT,
                        "agent not prepared for this batch size (hibf_gpu_hibf_membership_prepare)");
        unsigned const blocks =
            static_cast<unsigned>(std::min<uint64_t>(static_cast<uint64_t>(device_cus()) * 8u, (nq + 3u) / 4u));
        HIBF_TRY_HIP(launch_hibf_wave_membership(L, d_values, d_offsets, nq, d_thresholds, std::max(blocks, 1u), d_ids,
                                                 d_ids ? ids_cap : 0, d_hit_start, d_hit_count,
                                                 static_cast<uint64_t *>(a->overflow), s),
                     "membership launch");
        HIBF_TRY_HIP(launch_hibf_membership(L, d_values, d_offsets, nq, d_thresholds,
                                            static_cast<uint32_t *>(a->scratch), ob, d_ids, d_ids ? ids_cap : 0,
                                            d_hit_start, d_hit_count, static_cast<uint64_t *>(a->overflow), s),
                     "membership overflow launch");
        return HIBF_GPU_OK;
    }
    unsigned const blocks = hibf_blocks(nq, per_block, g_count_tuning.hibf_block_waves);
    if (a->scratch_bytes < static_cast<size_t>(blocks) * per_block)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT,
                    "agent not prepared for this batch size (hibf_gpu_hibf_membership_prepare)");
    HIBF_TRY_HIP(launch_hibf_membership(L, d_values, d_offsets, nq, d_thresholds, static_cast<uint32_t *>(a->scratch),
                                        blocks, d_ids, d_ids ? ids_cap : 0, d_hit_start, d_hit_count, nullptr, s),
                 "membership launch");
    return HIBF_GPU_OK;
}

} // namespace

int hibf_gpu_hibf_membership_prepare(hibf_gpu_hibf h, hibf_gpu_agent a, uint64_t max_queries)
{
    if (!h || !a)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (int rc = prepare_membership(h, a, std::max<uint64_t>(max_queries, 1)))
        return rc;
    HIBF_TRY_HIP(hipMemset(a->d_total + 5, 0, sizeof(unsigned long long)), "flag reset");
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_membership_launch(hibf_gpu_hibf h,
                                    hibf_gpu_agent a,
                                    uint64_t const * d_values,
                                    uint64_t const * d_offsets,
                                    uint64_t nq,
                                    uint16_t const * d_thresholds,
                                    uint64_t * d_ids,
                                    uint64_t ids_cap,
                                    uint64_t * d_hit_start,
                                    uint64_t * d_hit_count,
                                    void * stream)
{
    if (!h || !a || ((!d_offsets || !d_thresholds || !d_hit_start || !d_hit_count) && nq))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (nq == 0)
        return HIBF_GPU_OK;
    return enqueue_membership(h, a, d_values, d_offsets, nq, d_thresholds, d_ids, ids_cap, d_hit_start, d_hit_count,
                              static_cast<hipStream_t>(stream));
}

// Library-internal (serving loops), so C++ linkage inside this extern "C" block.
extern "C++" {
namespace hibf_amd
{
int membership_launch_capped(hibf_gpu_hibf h,
                             hibf_gpu_agent a,
                             uint64_t const * d_values,
                             uint64_t const * d_offsets,
                             uint64_t nq,
                             uint16_t const * d_thresholds,
                             uint64_t * d_ids,
                             uint32_t per_query_cap,
                             uint64_t * d_hit_start,
                             uint64_t * d_hit_count,
                             hipStream_t stream)
{
    if (!h || !a || ((!d_offsets || !d_thresholds || !d_hit_start || !d_hit_count || !d_ids) && nq))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (per_query_cap == 0u)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "per-query id cap must be > 0");
    if (nq == 0)
        return HIBF_GPU_OK;
    return enqueue_membership(h, a, d_values, d_offsets, nq, d_thresholds, d_ids, nq * per_query_cap, d_hit_start,
                              d_hit_count, stream, per_query_cap);
}
} // namespace hibf_amd
} // extern "C++"

int hibf_gpu_hibf_membership_launch_noalloc(hibf_gpu_hibf h,
                                            uint64_t const * d_values,
                                            uint64_t const * d_offsets,
                                            uint64_t nq,
                                            uint16_t const * d_thresholds,
                                            uint64_t * d_ids,
                                            uint64_t ids_cap,
                                            uint64_t * d_hit_start,
                                            uint64_t * d_hit_count,
                                            void * stream)
{
    if (!h || ((!d_offsets || !d_thresholds || !d_hit_start || !d_hit_count) && nq))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (nq == 0)
        return HIBF_GPU_OK;
    if (nq > kNoAllocMaxQueries)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT,
                    "hibf_gpu_hibf_membership_launch_noalloc: at most 65536 queries per call (got " +
                        std::to_string(nq) + ")");
    uint64_t const per_block = block_hit_list_bytes(h);
    uint64_t const max_blocks = (kNoAllocScratchBytes - kNoAllocHitListOffset) / per_block;
    if (max_blocks == 0)
        return fail(HIBF_GPU_ERR_UNSUPPORTED,
                    "hibf_gpu_hibf_membership_launch_noalloc: one query's hit list (" + std::to_string(per_block) +
                        " bytes) exceeds the 64 MiB workspace");
    unsigned char * ws = nullptr;
    if (int rc = noalloc_scratch(&ws))
        return rc;
    auto * const counters = reinterpret_cast<unsigned long long *>(ws);
    auto * const overflow = reinterpret_cast<uint64_t *>(ws + kNoAllocCounterBytes);
    auto * const hit_lists = reinterpret_cast<uint32_t *>(ws + kNoAllocHitListOffset);
    hipStream_t const s = static_cast<hipStream_t>(stream);
    HibfLaunch const L = launch_with(h, counters);
    // Every counter starts at zero: the previous call's cursor, totals and flags stay in the workspace.
    HIBF_TRY_HIP(launch_reset_counters(counters, 0xFFu, s), "counter reset");
    if (wave_kernel_fits(L))
    {
        unsigned const ob = static_cast<unsigned>(std::min<uint64_t>(overflow_blocks(h, nq), max_blocks));
        unsigned const blocks =
            static_cast<unsigned>(std::min<uint64_t>(static_cast<uint64_t>(device_cus()) * 8u, (nq + 3u) / 4u));
        HIBF_TRY_HIP(launch_hibf_wave_membership(L, d_values, d_offsets, nq, d_thresholds, std::max(blocks, 1u), d_ids,
                                                 d_ids ? ids_cap : 0, d_hit_start, d_hit_count, overflow, s),
                     "membership launch");
        HIBF_TRY_HIP(launch_hibf_membership(L, d_values, d_offsets, nq, d_thresholds, hit_lists, ob, d_ids,
                                            d_ids ? ids_cap : 0, d_hit_start, d_hit_count, overflow, s),
                     "membership overflow launch");
        return HIBF_GPU_OK;
    }
    unsigned const blocks = static_cast<unsigned>(
        std::min<uint64_t>(hibf_blocks(nq, per_block, g_count_tuning.hibf_block_waves), max_blocks));
    HIBF_TRY_HIP(launch_hibf_membership(L, d_values, d_offsets, nq, d_thresholds, hit_lists, blocks, d_ids,
                                        d_ids ? ids_cap : 0, d_hit_start, d_hit_count, nullptr, s),
                 "membership launch");
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_membership_for(hibf_gpu_hibf h,
                                 hibf_gpu_agent a,
                                 uint64_t const * d_values,
                                 uint64_t const * d_offsets,
                                 uint64_t nq,
                                 uint16_t const * d_thresholds,
                                 uint64_t * d_ids,
                                 uint64_t ids_cap,
                                 uint64_t * d_hit_offsets,
                                 uint64_t * h_total,
                                 void * stream)
{
    if (!h || !a || !h_total || ((!d_offsets || !d_thresholds || !d_hit_offsets) && nq))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t const s = static_cast<hipStream_t>(stream);
    *h_total = 0;
    if (nq == 0)
        return HIBF_GPU_OK;
    if (int rc = prepare_membership(h, a, nq))
        return rc;
    if (int rc = ensure(a->hit_start, a->hit_start_bytes, nq * 8u, "membership hits"))
        return rc;
    if (int rc = ensure(a->hit_count, a->hit_count_bytes, nq * 8u, "membership hits"))
        return rc;
    if (!a->ids_tmp)
        if (int rc = ensure(a->ids_tmp, a->ids_tmp_bytes, std::max<uint64_t>(nq, 1024u) * 8u, "membership ids"))
            return rc;
    for (int attempt = 0; attempt < 2; ++attempt)
    {
        uint64_t const cap = a->ids_tmp_bytes / 8u;
        // The dropped flag describes this call's result: a first attempt whose
        // temporary id buffer was too small must not leave it set.
        HIBF_TRY_HIP(launch_reset_counters(a->d_total, 1u << 5, s), "flag reset");
        if (int rc = enqueue_membership(h, a, d_values, d_offsets, nq, d_thresholds,
                                        static_cast<uint64_t *>(a->ids_tmp), cap,
                                        static_cast<uint64_t *>(a->hit_start), static_cast<uint64_t *>(a->hit_count),
                                        s))
            return rc;
        HIBF_TRY_HIP(hipMemcpyAsync(a->h_total, a->d_total, sizeof(unsigned long long), hipMemcpyDeviceToHost, s),
                     "total copy");
        HIBF_TRY_HIP(hipStreamSynchronize(s), "membership sync");
        uint64_t const total = *a->h_total;
        if (total <= cap)
            break;
        if (int rc = ensure(a->ids_tmp, a->ids_tmp_bytes, total * 8u, "membership ids"))
            return rc;
    }
    uint64_t const total = *a->h_total;
    *h_total = total;
    if (int rc = build_offsets(a, static_cast<uint64_t *>(a->hit_count), nq, nullptr, d_hit_offsets, s))
        return rc;
    if (total > ids_cap || (!d_ids && total))
    {
        HIBF_TRY_HIP(hipStreamSynchronize(s), "membership sync");
        return fail(HIBF_GPU_ERR_CAPACITY, "ids_cap too small; *h_total holds the needed size");
    }
    HIBF_TRY_HIP(launch_gather_hits(static_cast<uint64_t *>(a->ids_tmp), static_cast<uint64_t *>(a->hit_start),
                                    static_cast<uint64_t *>(a->hit_count), d_hit_offsets, nq, d_ids, s),
                 "gather launch");
    return HIBF_GPU_OK;
}

int hibf_gpu_hibf_bulk_count(hibf_gpu_hibf h,
                             hibf_gpu_agent a,
                             uint64_t const * d_values,
                             uint64_t const * d_offsets,
                             uint64_t nq,
                             uint64_t const * d_thresholds,
                             uint32_t counter_bytes,
                             void * d_counts,
                             void * stream)
{
    if (!h || !a || ((!d_offsets || !d_thresholds || !d_counts) && nq))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (counter_bytes != 1 && counter_bytes != 2 && counter_bytes != 4 && counter_bytes != 8)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "counter_bytes must be 1, 2, 4 or 8");
    if (h->user_bin_in_two_ibfs)
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "bulk_count on an HIBF whose tables list one user bin in two IBFs");
    if (nq == 0)
        return HIBF_GPU_OK;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    HibfLaunch const L = launch_of(h, a);
    HIBF_TRY_HIP(launch_reset_counters(a->d_total, 1u << 1, s), "counter reset");
    if (wave_kernel_fits(L))
    {
        unsigned const blocks =
            static_cast<unsigned>(std::min<uint64_t>(static_cast<uint64_t>(device_cus()) * 8u, (nq + 3u) / 4u));
        HIBF_TRY_HIP(launch_hibf_wave_count(L, d_values, d_offsets, nq, d_thresholds, counter_bytes, d_counts,
                                            std::max(blocks, 1u), s),
                     "hibf count launch");
        return HIBF_GPU_OK;
    }
    uint64_t const state_bytes = h->gstate_words * sizeof(uint32_t);
    unsigned const blocks = hibf_blocks(nq, state_bytes);
    uint32_t * scratch = nullptr;
    if (state_bytes)
    {
        if (int rc = ensure(a->scratch, a->scratch_bytes, static_cast<size_t>(blocks) * state_bytes, "count scratch"))
            return rc;
        scratch = static_cast<uint32_t *>(a->scratch);
    }
    HIBF_TRY_HIP(launch_hibf_count(L, d_values, d_offsets, nq, d_thresholds, counter_bytes, d_counts, scratch, blocks,
                                   s),
                 "hibf count launch");
    return HIBF_GPU_OK;
}

/* ------------------------------------------------------- HIBF build --- */

int hibf_gpu_config_default(hibf_gpu_config * c)
{
    if (!c)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null config");
    host::Config const d{};
    *c = hibf_gpu_config{d.number_of_user_bins, d.number_of_hash_functions, d.maximum_fpr, d.relaxed_fpr, d.threads,
                         d.sketch_bits, d.tmax, d.empty_bin_fraction, d.alpha, d.max_rearrangement_ratio,
                         d.disable_estimate_union, d.disable_rearrangement, d.track_occupancy, 0};
    return HIBF_GPU_OK;
}

namespace
{
host::Config to_host(hibf_gpu_config const & c)
{
    host::Config h{};
    h.number_of_user_bins = c.number_of_user_bins;
    h.number_of_hash_functions = c.number_of_hash_functions;
    h.maximum_fpr = c.maximum_fpr;
    h.relaxed_fpr = c.relaxed_fpr;
    h.threads = c.threads;
    h.sketch_bits = c.sketch_bits;
    h.tmax = c.tmax;
    h.empty_bin_fraction = c.empty_bin_fraction;
    h.alpha = c.alpha;
    h.max_rearrangement_ratio = c.max_rearrangement_ratio;
    h.disable_estimate_union = c.disable_estimate_union != 0;
    h.disable_rearrangement = c.disable_rearrangement != 0;
    h.track_occupancy = c.track_occupancy != 0;
    return h;
}

void from_host(host::Config const & h, hibf_gpu_config & c)
{
    c.tmax = h.tmax;
    c.track_occupancy = h.track_occupancy;
    c.disable_rearrangement = h.disable_rearrangement;
}

int copy_text(std::string const & t, char * out, uint64_t cap, uint64_t * needed)
{
    if (needed)
        *needed = t.size() + 1;
    if (!out || cap < t.size() + 1)
        return fail(HIBF_GPU_ERR_CAPACITY, "text buffer too small; *needed holds the size");
    std::memcpy(out, t.c_str(), t.size() + 1);
    return HIBF_GPU_OK;
}

// Device allocation freed when the owner goes out of scope.
struct DeviceBuffer
{
    void * p{nullptr};
    DeviceBuffer() = default;
    DeviceBuffer(DeviceBuffer const &) = delete;
    DeviceBuffer & operator=(DeviceBuffer const &) = delete;
    ~DeviceBuffer()
    {
        if (p)
            (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hibf_amd::device_malloc(&p, std::max<size_t>(bytes, 16u)); }
    uint8_t * u8() const { return static_cast<uint8_t *>(p); }
    uint64_t * u64() const { return static_cast<uint64_t *>(p); }
};

// User-bin cardinalities as the reference's build computes them for the
// layout: one HyperLogLog sketch of sketch_bits per user bin, each estimate
// truncated to an integer (src/hierarchical_interleaved_bloom_filter.cpp:
// 215-220; compute_sketches.cpp:27-43; estimate_kmer_counts.hpp).  Values come
// from the CSR on the device or, when d_values is null, from the synthetic
// stream of `seed` with h_sizes[u] values per user bin.
int estimate_kmer_counts(host::Config const & hc,
                         uint64_t const * d_values,
                         uint64_t const * d_offsets,
                         uint64_t const * h_sizes,
                         uint64_t seed,
                         std::vector<uint64_t> & counts,
                         DeviceBuffer & regs)
{
    uint64_t const n = hc.number_of_user_bins;
    uint32_t const bits = static_cast<uint32_t>(hc.sketch_bits);
    counts.assign(n, 0u);
    if (n > (~0ULL >> bits))
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "number_of_user_bins x 2^sketch_bits sketch registers exceed 2^64 bytes");
    HIBF_TRY_HIP(regs.alloc(static_cast<size_t>(n) << bits), "sketch registers");
    if (d_values)
    {
        if (int rc = hibf_gpu_hll_sketch(d_values, d_offsets, n, bits, regs.u8(), nullptr))
            return rc;
    }
    else if (int rc = hll_sketch_synthetic(seed, h_sizes, n, bits, regs.u8(), nullptr))
        return rc;
    std::vector<double> est(n);
    if (int rc = hll_estimates(regs.u8(), n, bits, est.data(), nullptr))
        return rc;
    for (uint64_t u = 0; u < n; ++u)
        counts[u] = static_cast<uint64_t>(est[u]);
    return HIBF_GPU_OK;
}

// Executes a build plan on the device: allocates the HIBF (zeroed arena) and
// runs the segment emplace over explicit values (device CSR) or the synthetic
// stream.
int run_plan(host::Plan const & plan,
             host::Layout const & layout,
             uint64_t n_user_bins,
             uint64_t const * d_ub_offsets,
             uint64_t const * d_values,
             uint64_t seed,
             double const * phase_s,
             hibf_gpu_hibf * out)
{
    auto const t_alloc = std::chrono::steady_clock::now();
    uint64_t const n = plan.ibfs.size();
    std::vector<hibf_gpu_ibf_geometry> geoms(n);
    std::vector<uint64_t> word_offsets(n), tb_offsets(n + 1, 0), next_ibf, user_bin;
    uint64_t words = 0;
    for (uint64_t i = 0; i < n; ++i)
    {
        host::PlannedIbf const & p = plan.ibfs[i];
        hibf_gpu_ibf_geometry g{};
        if (int rc = make_geometry(p.bins, p.bin_size, p.hash_funs, g))
            return rc;
        geoms[i] = g;
        word_offsets[i] = words;
        words += (g.n_words + 31u) / 32u * 32u; // 256-byte aligned IBFs
        tb_offsets[i + 1] = tb_offsets[i] + p.bins;
        next_ibf.insert(next_ibf.end(), p.next_ibf.begin(), p.next_ibf.end());
        user_bin.insert(user_bin.end(), p.user_bin.begin(), p.user_bin.end());
    }
    hibf_gpu_hibf h = nullptr;
    if (int rc = hibf_gpu_hibf_create(n, geoms.data(), nullptr, word_offsets.data(), tb_offsets.data(),
                                      next_ibf.data(), user_bin.data(), n_user_bins, &h))
        return rc;
    auto const t_fill = std::chrono::steady_clock::now();
    // Chunk the segments so that one block handles at most 4096 values.
    static_assert(sizeof(EmplaceSegment) == sizeof(host::Segment), "segment layouts differ");
    std::vector<EmplaceChunk> chunks;
    for (uint64_t s = 0; s < plan.segments.size(); ++s)
        for (uint64_t b = 0; b < plan.segments[s].count; b += 4096u)
            chunks.push_back({static_cast<uint32_t>(s), static_cast<uint32_t>(std::min<uint64_t>(4096u, plan.segments[s].count - b)), b});
    void * d_segs = nullptr;
    void * d_chunks = nullptr;
    void * d_chunk_ids = nullptr;
    auto cleanup = [&]()
    {
        for (void * p : {d_segs, d_chunks, d_chunk_ids})
            if (p)
                (void)hipFree(p);
    };
    hipError_t e = hibf_amd::device_malloc(&d_segs, std::max<size_t>(plan.segments.size(), 1) * sizeof(EmplaceSegment));
    if (e == hipSuccess)
        e = hipMemcpy(d_segs, plan.segments.data(), plan.segments.size() * sizeof(EmplaceSegment), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hibf_amd::device_malloc(&d_chunks, std::max<size_t>(chunks.size(), 1) * sizeof(EmplaceChunk));
    if (e == hipSuccess)
        e = hipMemcpy(d_chunks, chunks.data(), chunks.size() * sizeof(EmplaceChunk), hipMemcpyHostToDevice);
    if (e == hipSuccess && !plan.chunk_ids.empty())
    {
        e = hibf_amd::device_malloc(&d_chunk_ids, plan.chunk_ids.size());
        if (e == hipSuccess)
            e = hipMemcpy(d_chunk_ids, plan.chunk_ids.data(), plan.chunk_ids.size(), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = launch_segment_emplace(h->d_desc, h->d_arena, static_cast<EmplaceSegment *>(d_segs),
                                   static_cast<EmplaceChunk *>(d_chunks), chunks.size(), d_values, d_ub_offsets, seed,
                                   static_cast<uint8_t const *>(d_chunk_ids), plan.chunk_id_bytes, nullptr);
    if (e == hipSuccess)
        e = hipDeviceSynchronize();
    cleanup();
    if (e != hipSuccess)
    {
        hibf_gpu_hibf_destroy(h);
        return hip_fail(e, "hibf build");
    }
    auto const t_end = std::chrono::steady_clock::now();
    h->build_times[0] = phase_s[0];
    h->build_times[1] = phase_s[1];
    h->layout_times[0] = phase_s[2];
    h->layout_times[1] = phase_s[3];
    h->build_times[2] = std::chrono::duration<double>(t_fill - t_alloc).count();
    h->build_times[3] = std::chrono::duration<double>(t_end - t_fill).count();
    try
    {
        h->layout_text = layout.to_text();
    }
    catch (std::exception const &)
    {
        h->layout_text.clear();
    }
    *out = h;
    return HIBF_GPU_OK;
}

int get_layout(host::Config const & hc, std::vector<uint64_t> const & counts, char const * layout_text,
               host::Layout & layout)
{
    if (layout_text)
    {
        if (!host::Layout::from_text(layout_text, layout))
            return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "layout text could not be parsed");
        return HIBF_GPU_OK;
    }
    layout = host::compute_layout(hc, counts);
    return HIBF_GPU_OK;
}

// The layout over HyperLogLog sketches resident on the device, with union
// estimates and (unless disabled) similarity rearrangement, as the reference's
// compute_layout(config, kmer_counts, sketches): the union estimates of the DP
// and the clustering's merged-sketch estimates come from the device (sketches
// of at most 12 bits; wider ones are estimated on the host); the clustering's
// bookkeeping runs on host copies of the sketches.  times[0]
// and [1] get the union-estimation and rearrangement seconds.
int sketched_layout(host::Config const & hc,
                    std::vector<uint64_t> const & counts,
                    uint8_t const * d_regs,
                    host::Layout & layout,
                    double * times,
                    std::vector<uint64_t> const * positions = nullptr)
{
    uint64_t const n = hc.number_of_user_bins;
    uint32_t const bits = static_cast<uint32_t>(hc.sketch_bits);
    std::vector<uint8_t> regs(static_cast<size_t>(n) << bits);
    HIBF_TRY_HIP(hipDeviceSynchronize(), "sketch registers"); // the sketches may come from any stream
    HIBF_TRY_HIP(hipMemcpy(regs.data(), d_regs, regs.size(), hipMemcpyDeviceToHost), "sketch registers");
    std::vector<seqan::hibf::sketch::hyperloglog> sketches;
    sketches.reserve(n);
    for (uint64_t u = 0; u < n; ++u)
        sketches.push_back(seqan::hibf::sketch::hyperloglog::from_registers(static_cast<uint8_t>(bits),
                                                                            regs.data() + (u << bits)));
    std::unique_ptr<DeviceUnionEstimates> device_unions;
    std::unique_ptr<DevicePairUnions> device_pairs;
    host::Sketching sk;
    sk.sketches = &sketches;
    if (bits <= kUnionMaxBits)
    {
        if (int rc = DeviceUnionEstimates::create(d_regs, n, bits, device_unions))
            return rc;
        if (int rc = DevicePairUnions::create(bits, device_pairs))
            return rc;
        sk.unions = device_unions.get();
        sk.pair_unions = device_pairs.get();
    }
    try
    {
        layout = host::compute_layout(hc, counts, &sk, positions);
    }
    catch (std::out_of_range const &)
    {
        // An empty user bin: its distance to every other is 0 / 0 and the
        // clustering finds no closest pair (the reference fails the same way).
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "similarity rearrangement needs non-empty user bins");
    }
    catch (std::runtime_error const &)
    {
        return HIBF_GPU_ERR_DEVICE; // the union estimates set the message
    }
    times[0] = sk.union_seconds;
    times[1] = sk.rearrangement_seconds;
    return HIBF_GPU_OK;
}

// The layout of a build: the given text, or the DP over the user bins'
// HyperLogLog estimates (values from the device CSR, or the synthetic stream
// of `seed` when d_values is null), with union estimates from the sketches
// unless disable_estimate_union.  phase_s[0] .. [3] get the sketch, layout,
// union-estimation and rearrangement times.
int layout_for(host::Config const & hc,
               std::vector<uint64_t> const & sizes,
               char const * layout_text,
               uint64_t const * d_values,
               uint64_t const * d_offsets,
               uint64_t seed,
               host::Layout & layout,
               double * phase_s)
{
    using clock = std::chrono::steady_clock;
    if (layout_text)
    {
        auto const t0 = clock::now();
        int const rc = get_layout(hc, sizes, layout_text, layout);
        phase_s[1] = std::chrono::duration<double>(clock::now() - t0).count();
        return rc;
    }
    auto const t0 = clock::now();
    std::vector<uint64_t> kmer_counts;
    DeviceBuffer regs;
    if (int rc = estimate_kmer_counts(hc, d_values, d_offsets, d_values ? nullptr : sizes.data(), seed, kmer_counts,
                                      regs))
        return rc;
    auto const t1 = clock::now();
    int const rc = hc.disable_estimate_union ? get_layout(hc, kmer_counts, nullptr, layout)
                                             : sketched_layout(hc, kmer_counts, regs.u8(), layout, phase_s + 2);
    phase_s[0] = std::chrono::duration<double>(t1 - t0).count();
    phase_s[1] = std::chrono::duration<double>(clock::now() - t1).count();
    return rc;
}

} // namespace

int hibf_gpu_hibf_layout(hibf_gpu_hibf h, char * text, uint64_t cap, uint64_t * needed)
{
    if (!h || (!text && cap))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    return copy_text(h->layout_text, text, cap, needed);
}

uint64_t hibf_gpu_bin_size_in_bits(double fpr, uint64_t hash_count, uint64_t elements)
{
    return host::bin_size_in_bits(fpr, hash_count, elements);
}

int hibf_gpu_config_validate(hibf_gpu_config * c)
{
    if (!c)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null config");
    host::Config h = to_host(*c);
    std::string const e = host::validate_and_set_defaults(h);
    if (!e.empty())
        return fail(HIBF_GPU_ERR_CONFIG, e);
    from_host(h, *c);
    return HIBF_GPU_OK;
}

int hibf_gpu_layout_compute(hibf_gpu_config const * c,
                            uint64_t const * h_kmer_counts,
                            char * text,
                            uint64_t cap,
                            uint64_t * needed)
{
    if (!c || !h_kmer_counts)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    host::Config hc = to_host(*c);
    std::string const e = host::validate_and_set_defaults(hc);
    if (!e.empty())
        return fail(HIBF_GPU_ERR_CONFIG, e);
    std::vector<uint64_t> counts(h_kmer_counts, h_kmer_counts + hc.number_of_user_bins);
    try
    {
        return copy_text(host::compute_layout(hc, counts).to_text(), text, cap, needed);
    }
    catch (std::exception const & ex)
    {
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, ex.what());
    }
}

int hibf_gpu_layout_compute_sketched(hibf_gpu_config const * c,
                                     uint64_t const * h_kmer_counts,
                                     uint8_t const * d_registers,
                                     char * text,
                                     uint64_t cap,
                                     uint64_t * needed,
                                     double * h_times)
{
    if (!c || !h_kmer_counts || !d_registers)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    host::Config hc = to_host(*c);
    std::string const e = host::validate_and_set_defaults(hc);
    if (!e.empty())
        return fail(HIBF_GPU_ERR_CONFIG, e);
    std::vector<uint64_t> counts(h_kmer_counts, h_kmer_counts + hc.number_of_user_bins);
    try
    {
        host::Layout layout;
        double times[2]{};
        int const rc = hc.disable_estimate_union ? get_layout(hc, counts, nullptr, layout)
                                                 : sketched_layout(hc, counts, d_registers, layout, times);
        if (rc)
            return rc;
        if (h_times)
        {
            h_times[0] = times[0];
            h_times[1] = times[1];
        }
        return copy_text(layout.to_text(), text, cap, needed);
    }
    catch (std::exception const & ex)
    {
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, ex.what());
    }
}

int hibf_gpu_layout_compute_host_sketches(hibf_gpu_config const * c,
                                          uint64_t const * h_kmer_counts,
                                          uint8_t const * h_registers,
                                          uint64_t const * h_positions,
                                          uint64_t n_positions,
                                          char * text,
                                          uint64_t cap,
                                          uint64_t * needed,
                                          double * h_times)
{
    if (!c || !h_kmer_counts || !h_registers || (!h_positions && n_positions))
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    host::Config hc = to_host(*c);
    std::string const e = host::validate_and_set_defaults(hc);
    if (!e.empty())
        return fail(HIBF_GPU_ERR_CONFIG, e);
    uint64_t const n = hc.number_of_user_bins;
    uint32_t const bits = static_cast<uint32_t>(hc.sketch_bits);
    if (n > (~0ULL >> bits))
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "number_of_user_bins x 2^sketch_bits sketch registers exceed 2^64 bytes");
    std::vector<uint64_t> positions;
    if (h_positions)
    {
        if (n_positions == 0)
            return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "no user bins to lay out");
        positions.assign(h_positions, h_positions + n_positions);
        for (uint64_t const pos : positions)
            if (pos >= n)
                return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "position out of range");
    }
    try
    {
        std::vector<uint64_t> counts(h_kmer_counts, h_kmer_counts + n);
        host::Layout layout;
        double times[2]{};
        int rc = HIBF_GPU_OK;
        if (hc.disable_estimate_union)
            layout = host::compute_layout(hc, counts, nullptr, h_positions ? &positions : nullptr);
        else
        {
            DeviceBuffer d_regs;
            HIBF_TRY_HIP(d_regs.alloc(static_cast<size_t>(n) << bits), "sketch registers");
            HIBF_TRY_HIP(hipMemcpy(d_regs.p, h_registers, static_cast<size_t>(n) << bits, hipMemcpyHostToDevice),
                         "sketch registers");
            rc = sketched_layout(hc, counts, d_regs.u8(), layout, times, h_positions ? &positions : nullptr);
        }
        if (rc)
            return rc;
        if (h_times)
        {
            h_times[0] = times[0];
            h_times[1] = times[1];
        }
        return copy_text(layout.to_text(), text, cap, needed);
    }
    catch (std::exception const & ex)
    {
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, ex.what());
    }
}

namespace
{
// Host threads of the split-bin order replay (at most 16, the share of one GPU's host cores).
unsigned build_threads()
{
    unsigned const hw = std::thread::hardware_concurrency();
    return std::max(1u, std::min(16u, hw ? hw : 1u));
}

// The caller's user bins in their given order; tags index the sorted distinct copy on the device.
class ArraySource final : public host::ValueSource
{
public:
    ArraySource(uint64_t const * values, uint64_t const * offsets, std::vector<uint64_t> const & sorted,
                std::vector<uint64_t> const & sorted_offsets, uint64_t n_user_bins) :
        values_{values},
        offsets_{offsets},
        tags_(offsets[n_user_bins])
    {
        for (uint64_t u = 0; u < n_user_bins; ++u)
        {
            auto const b = sorted.begin() + static_cast<int64_t>(sorted_offsets[u]);
            auto const e = sorted.begin() + static_cast<int64_t>(sorted_offsets[u + 1]);
            for (uint64_t j = offsets[u]; j < offsets[u + 1]; ++j)
                tags_[j] = static_cast<uint32_t>(std::lower_bound(b, e, values[j]) - b);
        }
    }
    uint64_t read(uint32_t u, uint64_t pos, uint64_t cap, uint64_t * vals, uint32_t * tags) const override
    {
        uint64_t const n = offsets_[u + 1] - offsets_[u];
        if (pos >= n)
            return 0;
        uint64_t const k = std::min(cap, n - pos);
        std::memcpy(vals, values_ + offsets_[u] + pos, k * sizeof(uint64_t));
        std::memcpy(tags, tags_.data() + offsets_[u] + pos, k * sizeof(uint32_t));
        return k;
    }

private:
    uint64_t const * values_;
    uint64_t const * offsets_;
    std::vector<uint32_t> tags_;
};

// The synthetic user bins: value i of user bin u is hibf_gpu_synth_value(seed, u, i).
class SynthSource final : public host::ValueSource
{
public:
    SynthSource(uint64_t seed, std::vector<uint64_t> const & sizes) : seed_{seed}, sizes_{sizes} {}
    uint64_t read(uint32_t u, uint64_t pos, uint64_t cap, uint64_t * vals, uint32_t * tags) const override
    {
        uint64_t const n = sizes_[u];
        if (pos >= n)
            return 0;
        uint64_t const k = std::min(cap, n - pos);
        for (uint64_t j = 0; j < k; ++j)
        {
            vals[j] = hibf_gpu_synth_value(seed_, u, pos + j);
            tags[j] = static_cast<uint32_t>(pos + j);
        }
        return k;
    }

private:
    uint64_t seed_;
    std::vector<uint64_t> const & sizes_;
};
} // namespace

int hibf_gpu_hibf_build(hibf_gpu_config const * c,
                        uint64_t const * h_ub_offsets,
                        uint64_t const * h_values,
                        char const * layout_text,
                        hibf_gpu_hibf * out)
{
    if (!c || !h_ub_offsets || !h_values || !out)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    host::Config hc = to_host(*c);
    std::string const e = host::validate_and_set_defaults(hc);
    if (!e.empty())
        return fail(HIBF_GPU_ERR_CONFIG, e);
    uint64_t const nub = hc.number_of_user_bins;
    if (nub >= (1ULL << 32))
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "too many user bins");
    try
    {
        // Each user bin is a set (the reference inserts into a hash set): sort + unique.
        std::vector<uint64_t> offs(nub + 1, 0);
        std::vector<uint64_t> vals;
        vals.reserve(h_ub_offsets[nub]);
        for (uint64_t u = 0; u < nub; ++u)
        {
            if (h_ub_offsets[u + 1] < h_ub_offsets[u])
                return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "user-bin offsets must not decrease");
            size_t const b = vals.size();
            vals.insert(vals.end(), h_values + h_ub_offsets[u], h_values + h_ub_offsets[u + 1]);
            std::sort(vals.begin() + b, vals.end());
            vals.erase(std::unique(vals.begin() + b, vals.end()), vals.end());
            offs[u + 1] = vals.size();
        }
        host::BuildInput in;
        in.ub_size.resize(nub);
        for (uint64_t u = 0; u < nub; ++u)
            in.ub_size[u] = offs[u + 1] - offs[u];
        in.union_size = [&](std::vector<uint32_t> const & ubs)
        {
            std::vector<uint64_t> all;
            for (uint32_t u : ubs)
                all.insert(all.end(), vals.begin() + offs[u], vals.begin() + offs[u + 1]);
            std::sort(all.begin(), all.end());
            return static_cast<uint64_t>(std::unique(all.begin(), all.end()) - all.begin());
        };
        DeviceBuffer d_vals, d_offs;
        HIBF_TRY_HIP(d_vals.alloc(vals.size() * 8u), "build values");
        HIBF_TRY_HIP(hipMemcpy(d_vals.p, vals.data(), vals.size() * 8u, hipMemcpyHostToDevice), "build values");
        HIBF_TRY_HIP(d_offs.alloc(offs.size() * 8u), "build offsets");
        HIBF_TRY_HIP(hipMemcpy(d_offs.p, offs.data(), offs.size() * 8u, hipMemcpyHostToDevice), "build offsets");
        // The layout sees HyperLogLog estimates, as in the reference; IBF sizes use the exact sets
        // (the reference sizes each IBF from the k-mers it inserts).
        double phase_s[4]{};
        host::Layout layout;
        if (int rc = layout_for(hc, in.ub_size, layout_text, d_vals.u64(), d_offs.u64(), 0, layout, phase_s))
            return rc;
        std::unique_ptr<ArraySource> source;
        if (g_count_tuning.split_order)
        {
            for (uint64_t u = 0; u < nub; ++u)
                if (in.ub_size[u] >= (1ULL << 32))
                    return fail(HIBF_GPU_ERR_UNSUPPORTED, "a user bin of 2^32 distinct values or more");
            source = std::make_unique<ArraySource>(h_values, h_ub_offsets, vals, offs, nub);
            in.values = source.get();
            in.threads = build_threads();
        }
        host::Plan plan;
        std::string const pe = host::build_plan(hc, layout, in, plan);
        if (!pe.empty())
            return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, pe);
        return run_plan(plan, layout, nub, d_offs.u64(), d_vals.u64(), 0, phase_s, out);
    }
    catch (std::exception const & ex)
    {
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, ex.what());
    }
}

int hibf_gpu_hibf_plan_stats(hibf_gpu_config const * c,
                             uint64_t const * h_ub_sizes,
                             char const * layout_text,
                             uint64_t * n_ibfs,
                             uint64_t * n_words,
                             uint64_t * n_segments,
                             uint64_t * n_emplaces)
{
    if (!c || !h_ub_sizes)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    host::Config hc = to_host(*c);
    std::string const e = host::validate_and_set_defaults(hc);
    if (!e.empty())
        return fail(HIBF_GPU_ERR_CONFIG, e);
    try
    {
        host::BuildInput in;
        in.ub_size.assign(h_ub_sizes, h_ub_sizes + hc.number_of_user_bins);
        host::Layout layout;
        if (int rc = get_layout(hc, in.ub_size, layout_text, layout))
            return rc;
        host::Plan plan;
        std::string const pe = host::build_plan(hc, layout, in, plan);
        if (!pe.empty())
            return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, pe);
        uint64_t words = 0, emplaces = 0;
        for (auto const & p : plan.ibfs)
        {
            hibf_gpu_ibf_geometry g{};
            if (int rc = make_geometry(p.bins, p.bin_size, p.hash_funs, g))
                return rc;
            words += (g.n_words + 31u) / 32u * 32u;
        }
        for (auto const & s : plan.segments)
            emplaces += s.count;
        if (n_ibfs)
            *n_ibfs = plan.ibfs.size();
        if (n_words)
            *n_words = words;
        if (n_segments)
            *n_segments = plan.segments.size();
        if (n_emplaces)
            *n_emplaces = emplaces;
        return HIBF_GPU_OK;
    }
    catch (std::exception const & ex)
    {
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, ex.what());
    }
}

int hibf_gpu_hibf_build_synthetic(hibf_gpu_config const * c,
                                  uint64_t const * h_ub_sizes,
                                  uint64_t seed,
                                  char const * layout_text,
                                  hibf_gpu_hibf * out)
{
    if (!c || !h_ub_sizes || !out)
        return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    host::Config hc = to_host(*c);
    std::string const e = host::validate_and_set_defaults(hc);
    if (!e.empty())
        return fail(HIBF_GPU_ERR_CONFIG, e);
    uint64_t const nub = hc.number_of_user_bins;
    if (nub >= (1ULL << 28))
        return fail(HIBF_GPU_ERR_UNSUPPORTED, "the synthetic stream supports < 2^28 user bins");
    try
    {
        host::BuildInput in;
        in.ub_size.assign(h_ub_sizes, h_ub_sizes + nub);
        for (uint64_t s : in.ub_size)
            if (s >= (1ULL << 36))
                return fail(HIBF_GPU_ERR_UNSUPPORTED, "the synthetic stream supports < 2^36 values per user bin");
        double phase_s[4]{};
        host::Layout layout;
        if (int rc = layout_for(hc, in.ub_size, layout_text, nullptr, nullptr, seed, layout, phase_s))
            return rc;
        SynthSource const source{seed, in.ub_size};
        if (g_count_tuning.split_order)
        {
            for (uint64_t u = 0; u < nub; ++u)
                if (in.ub_size[u] >= (1ULL << 32))
                    return fail(HIBF_GPU_ERR_UNSUPPORTED, "a user bin of 2^32 values or more");
            in.values = &source;
            in.threads = build_threads();
        }
        host::Plan plan;
        std::string const pe = host::build_plan(hc, layout, in, plan);
        if (!pe.empty())
            return fail(HIBF_GPU_ERR_INVALID_ARGUMENT, pe);
        return run_plan(plan, layout, nub, nullptr, nullptr, seed, phase_s, out);
    }
    catch (std::exception const & ex)
    {
        return fail(HIBF_GPU_ERR_OUT_OF_MEMORY, ex.what());
    }
}

} // extern "C"
