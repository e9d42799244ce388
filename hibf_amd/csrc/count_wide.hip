// bulk_count for wide IBFs (bin_words even and >= 64, e.g. 8,192 technical
// bins = 128 words = 1 KiB rows): the HBM-bound shape of the flat-IBF workload.
//
// One block (4 waves) per query.  Each lane owns two adjacent words of a
// 128-word tile and fetches them with one 16-byte load per row, so one wave
// instruction moves a whole 1 KiB row.  The row starts of a value are
// wave-uniform (readlane of the lane that hashed it), U values are in flight
// per wave before their hits are counted.  Counters are CB-bit fields packed in
// 32-bit LDS words: 16-bit pairs (16 KiB per 8,192 bins) or 8-bit quads
// (8 KiB, which lets more small blocks share a CU).  A query longer than
// 2^CB - 1 values is processed in segments whose counters are flushed (added)
// to the output, so no packed counter can carry into its neighbour; the sum of
// the segments modulo 2^(8 * counter_bytes) is the reference's wrapped count.
#include "device_common.hpp"
#include "kernels.hpp"

namespace hibf_amd
{

namespace
{

constexpr uint32_t kWideTileWords = 128; // 8,192 bins per tile

template <typename T>
__device__ __forceinline__ void put(void * out, uint64_t i, uint32_t v, bool accumulate)
{
    T * p = static_cast<T *>(out) + i;
    *p = accumulate ? static_cast<T>(*p + static_cast<T>(v)) : static_cast<T>(v);
}

__device__ __forceinline__ void put_counter(void * out, uint64_t i, uint32_t cb, uint32_t v, bool accumulate)
{
    switch (cb)
    {
    case 1: put<uint8_t>(out, i, v, accumulate); break;
    case 2: put<uint16_t>(out, i, v, accumulate); break;
    case 4: put<uint32_t>(out, i, v, accumulate); break;
    default: put<uint64_t>(out, i, v, accumulate); break;
    }
}

// Writes nb counters of a tile (CB-bit fields packed in lds) to out[first ..).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));

// A lane's two words of a row (16 bytes); NT: non-temporal (streaming) load.
template <bool NT>
__device__ __forceinline__ ulonglong2 load_pair(uint64_t const * p)
{
    if constexpr (NT)
    {
        u64x2 const v = __builtin_nontemporal_load(reinterpret_cast<u64x2 const *>(p));
        return make_ulonglong2(v.x, v.y);
    }
    else
        return *reinterpret_cast<ulonglong2 const *>(p);
}

template <int CB>
__device__ void store_tile(uint32_t const * lds,
                           uint32_t nb,
                           void * counts,
                           uint64_t first,
                           uint32_t cb,
                           bool accumulate,
                           bool nt,
                           int tid,
                           int nthreads)
{
    constexpr uint32_t per_word = 32u / CB;
    constexpr uint32_t field = (1u << CB) - 1u;
    if (!accumulate && cb == 2u && (first % 8u) == 0u)
    {
        // 8 uint16 counters per 16-byte store.
        uint32_t const n8 = nb / 8u;
        uint4 * const out = reinterpret_cast<uint4 *>(static_cast<uint16_t *>(counts) + first);
        for (uint32_t i = static_cast<uint32_t>(tid); i < n8; i += static_cast<uint32_t>(nthreads))
        {
            uint4 w;
            if constexpr (CB == 16)
            {
                w = *reinterpret_cast<uint4 const *>(lds + 4u * i);
            }
            else
            {
                uint32_t const a = lds[2u * i], b = lds[2u * i + 1u];
                w.x = (a & 0xFFu) | ((a & 0xFF00u) << 8);
                w.y = ((a >> 16) & 0xFFu) | ((a >> 8) & 0xFF0000u);
                w.z = (b & 0xFFu) | ((b & 0xFF00u) << 8);
                w.w = ((b >> 16) & 0xFFu) | ((b >> 8) & 0xFF0000u);
            }
            if (nt) // streaming store: the counters are not read again by this kernel
                __builtin_nontemporal_store(u32x4{w.x, w.y, w.z, w.w}, reinterpret_cast<u32x4 *>(out) + i);
            else
                out[i] = w;
        }
        for (uint32_t i = n8 * 8u + static_cast<uint32_t>(tid); i < nb; i += static_cast<uint32_t>(nthreads))
            static_cast<uint16_t *>(counts)[first + i] =
                static_cast<uint16_t>((lds[i / per_word] >> (CB * (i % per_word))) & field);
        return;
    }
    for (uint32_t i = static_cast<uint32_t>(tid); i < nb; i += static_cast<uint32_t>(nthreads))
        put_counter(counts, first + i, cb, (lds[i / per_word] >> (CB * (i % per_word))) & field, accumulate);
}

// HF > 0 fixes the number of hash functions (== g.hash_funs) at compile time,
// so only HF row starts per value occupy registers; HF == 0 reads it at run time.
// CB is the LDS counter width in bits (8 or 16).  Instances that fit 64 VGPRs
// are held to them (8 waves per SIMD); 8-bit counters at U = 4 need more (at
// 64 they spilled 20 bytes per lane), so they get the registers instead.
// NT: non-temporal row loads.
template <int U, int HF, int CB, bool NT = false>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu((HF > 0 && (U < 4 || (U == 4 && CB == 16))) ? 8 : 1, 8))) count_wide_kernel(IbfView g,
                                                         uint64_t const * __restrict__ values,
                                                         uint64_t const * __restrict__ offsets,
                                                         uint64_t nq,
                                                         uint32_t counter_bytes,
                                                         uint32_t nt_stores,
                                                         void * __restrict__ counts)
{
    extern __shared__ uint32_t lds[]; // kWideTileWords * 64 counters of CB bits
    constexpr uint32_t per_word = 32u / CB;
    constexpr uint32_t segment_values = (1u << CB) - 1u; // a packed counter never overflows
    int const tid = static_cast<int>(threadIdx.x);
    int const nthreads = static_cast<int>(blockDim.x);
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const nwaves = nthreads >> 6;
    uint64_t const bw = g.bin_words;
    uint32_t const hf = g.hash_funs;
    uint64_t const * __restrict__ words = g.words;
    constexpr int KMAX = HF > 0 ? HF : 5;

    for (uint64_t q = blockIdx.x; q < nq; q += gridDim.x)
    {
        uint64_t const beg = offsets[q];
        uint32_t const L = static_cast<uint32_t>(offsets[q + 1] - beg);
        uint64_t const * __restrict__ vals = values + beg;
        for (uint32_t w0 = 0; w0 < g.bin_words; w0 += kWideTileWords)
        {
            uint32_t const tw = min(kWideTileWords, g.bin_words - w0);
            uint32_t const wl = 2u * static_cast<uint32_t>(lane); // first of my two words in the tile
            bool const mine = wl < tw;                          // tw is even: both words or none
            uint64_t const m0 = mine ? tail_mask(g, w0 + wl) : 0ULL;
            uint64_t const m1 = mine ? tail_mask(g, w0 + wl + 1u) : 0ULL;
            uint32_t seg0 = 0;
            do
            {
                uint32_t const segL = min(segment_values, L - seg0);
                for (uint32_t i = static_cast<uint32_t>(tid); i < tw * 64u / per_word;
                     i += static_cast<uint32_t>(nthreads))
                    lds[i] = 0u;
                __syncthreads();
                for (uint32_t base = static_cast<uint32_t>(wave) * 64u; base < segL;
                     base += static_cast<uint32_t>(nwaves) * 64u)
                {
                    uint32_t const nv = min(64u, segL - base);
                    uint64_t const v = static_cast<uint32_t>(lane) < nv ? vals[seg0 + base + lane] : 0ULL;
                    uint64_t rows[KMAX];
#pragma unroll
                    for (int k = 0; k < KMAX; ++k)
                        rows[k] = (HF > 0 || static_cast<uint32_t>(k) < hf)
                                      ? row_word(v, hash_seed(k), g.hash_shift, g.bin_size, bw) + w0
                                      : 0ULL;
                    for (uint32_t s = 0; s < nv; s += U)
                    {
                        uint64_t x0[U], x1[U];
#pragma unroll
                        for (int u = 0; u < U; ++u)
                        {
                            x0[u] = 0ULL;
                            x1[u] = 0ULL;
                            uint32_t const idx = s + static_cast<uint32_t>(u);
                            if (idx < nv)
                            {
                                // Row start of value idx (uniform) + my word offset.
                                uint64_t const r0 = readlane64(rows[0], static_cast<int>(idx)) + wl;
                                if (mine)
                                {
                                    ulonglong2 a = load_pair<NT>(words + r0);
#pragma unroll
                                    for (int k = 1; k < KMAX; ++k)
                                        if (HF > 0 || static_cast<uint32_t>(k) < hf)
                                        {
                                            uint64_t const rk = readlane64(rows[k], static_cast<int>(idx)) + wl;
                                            ulonglong2 const b = load_pair<NT>(words + rk);
                                            a.x &= b.x;
                                            a.y &= b.y;
                                        }
                                    x0[u] = a.x & m0;
                                    x1[u] = a.y & m1;
                                }
                            }
                        }
#pragma unroll
                        for (int u = 0; u < U; ++u)
                        {
                            uint64_t bits = x0[u];
                            while (bits)
                            {
                                uint32_t const b = wl * 64u + static_cast<uint32_t>(__builtin_ctzll(bits));
                                atomicAdd(&lds[b / per_word], 1u << (CB * (b % per_word)));
                                bits &= bits - 1ULL;
                            }
                            bits = x1[u];
                            while (bits)
                            {
                                uint32_t const b = (wl + 1u) * 64u + static_cast<uint32_t>(__builtin_ctzll(bits));
                                atomicAdd(&lds[b / per_word], 1u << (CB * (b % per_word)));
                                bits &= bits - 1ULL;
                            }
                        }
                    }
                }
                __syncthreads();
                uint32_t const first_bin = w0 * 64u;
                uint32_t const nb = min(tw * 64u, g.bins - first_bin);
                store_tile<CB>(lds, nb, counts, q * g.bins + first_bin, counter_bytes, seg0 != 0, nt_stores != 0u, tid,
                               nthreads);
                __syncthreads();
                seg0 += segL;
            } while (seg0 < L);
        }
    }
}

} // namespace

CountTuning g_count_tuning{};
uint64_t g_alloc_max_bytes = 0;

bool count_wide_applies(IbfView const & g)
{
    return g.bin_words >= 64u && (g.bin_words % 2u) == 0u && g_count_tuning.wide_kernel != 0u;
}

hipError_t launch_count_wide(IbfView const & g,
                             uint64_t const * values,
                             uint64_t const * offsets,
                             uint64_t nq,
                             uint32_t counter_bytes,
                             void * counts,
                             unsigned max_blocks,
                             hipStream_t stream)
{
    if (nq == 0)
        return hipSuccess;
    bool const narrow = g_count_tuning.counter_bits == 8u;
    size_t const lds = static_cast<size_t>(kWideTileWords) * (narrow ? 16u : 32u) * sizeof(uint32_t);
    unsigned const grid = static_cast<unsigned>(nq < max_blocks ? nq : max_blocks);
    dim3 const blk(64u * g_count_tuning.wide_waves);
    using kernel_t = void (*)(IbfView, uint64_t const *, uint64_t const *, uint64_t, uint32_t, uint32_t, void *);
    // Non-temporal row loads, for 2 or 3 hash functions: [counter bits 16, 8][unroll 1, 2, 4, 8][2, 3]
    static kernel_t const table_nt[2][4][2] = {
        {{count_wide_kernel<1, 2, 16, true>, count_wide_kernel<1, 3, 16, true>},
         {count_wide_kernel<2, 2, 16, true>, count_wide_kernel<2, 3, 16, true>},
         {count_wide_kernel<4, 2, 16, true>, count_wide_kernel<4, 3, 16, true>},
         {count_wide_kernel<8, 2, 16, true>, count_wide_kernel<8, 3, 16, true>}},
        {{count_wide_kernel<1, 2, 8, true>, count_wide_kernel<1, 3, 8, true>},
         {count_wide_kernel<2, 2, 8, true>, count_wide_kernel<2, 3, 8, true>},
         {count_wide_kernel<4, 2, 8, true>, count_wide_kernel<4, 3, 8, true>},
         {count_wide_kernel<8, 2, 8, true>, count_wide_kernel<8, 3, 8, true>}},
    };
    // [counter bits 16, 8][unroll 1, 2, 4, 8][hash functions: run time, 2, 3]
    static kernel_t const table[2][4][3] = {
        {{count_wide_kernel<1, 0, 16>, count_wide_kernel<1, 2, 16>, count_wide_kernel<1, 3, 16>},
         {count_wide_kernel<2, 0, 16>, count_wide_kernel<2, 2, 16>, count_wide_kernel<2, 3, 16>},
         {count_wide_kernel<4, 0, 16>, count_wide_kernel<4, 2, 16>, count_wide_kernel<4, 3, 16>},
         {count_wide_kernel<8, 0, 16>, count_wide_kernel<8, 2, 16>, count_wide_kernel<8, 3, 16>}},
        {{count_wide_kernel<1, 0, 8>, count_wide_kernel<1, 2, 8>, count_wide_kernel<1, 3, 8>},
         {count_wide_kernel<2, 0, 8>, count_wide_kernel<2, 2, 8>, count_wide_kernel<2, 3, 8>},
         {count_wide_kernel<4, 0, 8>, count_wide_kernel<4, 2, 8>, count_wide_kernel<4, 3, 8>},
         {count_wide_kernel<8, 0, 8>, count_wide_kernel<8, 2, 8>, count_wide_kernel<8, 3, 8>}},
    };
    uint32_t const u = g_count_tuning.wide_unroll;
    int const ui = u == 1u ? 0 : u == 2u ? 1 : u == 8u ? 3 : 2;
    int const hi = g.hash_funs == 2u ? 1 : g.hash_funs == 3u ? 2 : 0;
    kernel_t const k = g_count_tuning.wide_nt_loads && hi != 0 ? table_nt[narrow ? 1 : 0][ui][hi - 1]
                                                              : table[narrow ? 1 : 0][ui][hi];
    hipLaunchKernelGGL(k, dim3(grid), blk, lds, stream, g, values, offsets, nq, counter_bytes,
                       g_count_tuning.nt_stores, counts);
    return hipGetLastError();
}

} // namespace hibf_amd
