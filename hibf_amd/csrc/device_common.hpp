// Device-side building blocks shared by the HIBF query kernels (gfx950).
//
// Bit-matrix layout (interleaved_bloom_filter.hpp:110-126, bit_vector.hpp:1045-1054):
// bit (row r, technical bin b) lives at bit position r*technical_bins + b, i.e.
// word r*bin_words + b/64, bit b%64, LSB first.  A "row" is the bin_words
// consecutive words that one hash position selects across all technical bins.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hibf_amd
{

// interleaved_bloom_filter.hpp:172-176
__device__ __forceinline__ uint64_t hash_seed(int k)
{
    switch (k)
    {
    case 0: return 13572355802537770549ULL;
    case 1: return 13043817825332782213ULL;
    case 2: return 10650232656628343401ULL;
    case 3: return 16499269484942379435ULL;
    default: return 4893150838803335377ULL;
    }
}

// First word of the row selected by value v under seed k.
// interleaved_bloom_filter.hpp:185-195 returns the bit position
// fastrange(...) * technical_bins; technical_bins = 64 * bin_words and the
// product never exceeds the matrix size, so position / 64 == fastrange * bin_words.
__device__ __forceinline__ uint64_t row_word(uint64_t v, uint64_t seed, uint32_t hash_shift, uint64_t bin_size,
                                             uint64_t bin_words)
{
    uint64_t h = v * seed;
    h ^= h >> hash_shift;
    h *= 11400714819323198485ULL;
    return __umul64hi(h, bin_size) * bin_words;
}

// Synthetic value stream used by generator-driven bulk emplace: value i of
// user bin (or technical bin) `bin` under `seed`.  fmix64 (the MurmurHash3
// finaliser) is a bijection of 2^64, and (bin << 36 | i) + seed * golden is
// injective for i < 2^36, bin < 2^28, so every (bin, i) pair of one seed gives
// a distinct, uniformly spread value.
__host__ __device__ __forceinline__ uint64_t synth_value(uint64_t seed, uint64_t bin, uint64_t i)
{
    uint64_t k = ((bin << 36) | i) + seed * 0x9E3779B97F4A7C15ULL;
    k ^= k >> 33;
    k *= 0xFF51AFD7ED558CCDULL;
    k ^= k >> 33;
    k *= 0xC4CEB9FE1A85EC53ULL;
    k ^= k >> 33;
    return k;
}

// Geometry of one IBF as the kernels need it.
struct IbfView
{
    uint64_t const * words; // bit matrix (device)
    uint64_t bin_size;      // rows
    uint32_t bins;          // bin_count()
    uint32_t bin_words;     // ceil(bins / 64)
    uint32_t hash_shift;    // countl_zero(bin_size)
    uint32_t hash_funs;     // 1..5
};

__device__ __forceinline__ uint64_t tail_mask(IbfView const & g, uint32_t w)
{
    uint32_t const r = g.bins & 63u;
    return (w + 1u == g.bin_words && r) ? ((1ULL << r) - 1ULL) : ~0ULL;
}

// Row-word load; NT marks it non-temporal (streaming) so the line is not kept
// in the caches for reuse that random rows never get.
template <bool NT>
__device__ __forceinline__ uint64_t load_row_word(uint64_t const * p)
{
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t x, int src)
{
    uint32_t lo = static_cast<uint32_t>(x);
    uint32_t hi = static_cast<uint32_t>(x >> 32);
    lo = __shfl(lo, src, 64);
    hi = __shfl(hi, src, 64);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Of the lanes with `store` set, keeps the highest lane of each id and clears
// the others (a wave-uniform loop over the storing lanes; nothing to do for
// fewer than two).  Used where one instruction would store several lanes'
// values to the same address and the last in bin order must win.
__device__ __forceinline__ bool highest_lane_per_id(bool store, uint32_t id, int lane)
{
    uint64_t rest = __ballot(store);
    if (__popcll(rest) < 2)
        return store;
    while (rest)
    {
        int const j = __builtin_ctzll(rest);
        rest &= rest - 1ULL;
        uint32_t const idj = static_cast<uint32_t>(__shfl(static_cast<int>(id), j, 64));
        if (store && lane < j && id == idj)
            store = false;
    }
    return store;
}

// Inclusive prefix sum of a uint32 over the 64 lanes of a wave: lane i gets
// x[0] + ... + x[i] modulo 2^32.  DPP row shifts by 1, 2, 4 and 8 lanes build
// the prefix inside each row of 16 lanes (lanes whose source lies before the
// row start add 0); broadcasting lane 15 of a row into the next (rows 1, 3)
// and lane 31 into rows 2 and 3 adds the totals of the rows before.
__device__ __forceinline__ uint32_t wave_inclusive_add_u32(uint32_t x)
{
    int v = static_cast<int>(x);
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false); // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false); // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false); // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false); // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false); // row_bcast:15, rows 1 and 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false); // row_bcast:31, rows 2 and 3
    return static_cast<uint32_t>(v);
}

// Values every lane of the wave holds alike, moved to scalar registers, so
// that addresses built from them are uniform.
__device__ __forceinline__ uint32_t uniform32(uint32_t x)
{
    return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x)));
}
__device__ __forceinline__ uint64_t uniform64(uint64_t x)
{
    uint32_t const lo = uniform32(static_cast<uint32_t>(x));
    uint32_t const hi = uniform32(static_cast<uint32_t>(x >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Load of a 32- or 64-bit word of memory that no thread writes while the
// kernel runs: through the constant address space, so that a uniform address
// becomes a scalar load served by the scalar data cache.
template <typename T>
__device__ __forceinline__ T const_load(T const * p)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "scalar loads are of dwords");
    using const_ptr = T const __attribute__((address_space(4))) *;
    return *(const_ptr)p; // address-space cast: flat and constant addresses of global memory are the same
}

__device__ __forceinline__ uint64_t readlane64(uint64_t x, int src)
{
    uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x)), src);
    uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x >> 32)), src);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Adds, for every value of vals[0..L) and every technical bin b of the word
// tile [w0, w0+tw), one to lds[(word - w0) * 64 + bit] when all h rows have the
// bin's bit set (bulk_contains + counting_vector += in one pass,
// src/interleaved_bloom_filter.cpp:204-291, counting_vector.hpp:279-311).
//
// A wave takes 64 values at a time; lane i hashes value i.  LPV lanes then
// share one value (VPS = 64 / LPV values per load step); each lane owns CW
// words of the tile: w0 + (lane % LPV) + c * LPV.  Counters are 32-bit LDS
// words updated with ds_add_u32 (values of one step may hit the same bin).
// Callers truncate them to the agent's counter width, which gives the
// reference's modulo-2^w wrap exactly.
//
// HF > 0 fixes the number of hash functions at compile time (it must equal
// g.hash_funs), which keeps only HF row starts per value in registers; HF == 0
// handles any count up to 5 at run time.
//
// With 2 or 4 lanes per value, a lane sees LPV values of each 64-value chunk,
// all for the same tile words.  Their masks are summed in bit-sliced registers
// (plane p holds bit p of each bin's count) and added to LDS once per chunk:
// one atomic per distinct bin instead of one per set bit and value, which
// matters because most values of a query set the same bin.
template <int LPV, int CW, int U, bool NT = false, int HF = 0, bool SLICE = true>
__device__ __forceinline__ void count_into_lds(IbfView const & g,
                                               uint64_t const * __restrict__ vals,
                                               uint32_t L,
                                               uint32_t w0,
                                               uint32_t tw,
                                               uint32_t * lds,
                                               int wave,
                                               int nwaves,
                                               int lane)
{
    constexpr int VPS = 64 / LPV;
    constexpr int KMAX = HF > 0 ? HF : 5;
    constexpr bool SLICED = SLICE && LPV >= 2 && LPV <= 4;
    constexpr int PLANES = LPV >= 4 ? 3 : 2; // counts up to LPV per chunk
    int const sub = lane / LPV;
    uint32_t const wl = static_cast<uint32_t>(lane % LPV);
    uint64_t const bw = g.bin_words;
    uint32_t const hf = g.hash_funs;
    uint64_t const * __restrict__ words = g.words;

    for (uint32_t base = static_cast<uint32_t>(wave) * 64u; base < L; base += static_cast<uint32_t>(nwaves) * 64u)
    {
        uint32_t const nv = min(64u, L - base);
        uint64_t const v = static_cast<uint32_t>(lane) < nv ? vals[base + lane] : 0ULL;
        uint64_t rows[KMAX];
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
            rows[k] = (HF > 0 || static_cast<uint32_t>(k) < hf) ? row_word(v, hash_seed(k), g.hash_shift, g.bin_size, bw)
                                                                 : 0ULL;

        uint64_t plane[PLANES][CW];
#pragma unroll
        for (int p = 0; p < PLANES; ++p)
#pragma unroll
            for (int c = 0; c < CW; ++c)
                plane[p][c] = 0ULL;

        for (uint32_t s = 0; s < nv; s += VPS * U)
        {
            uint64_t x[U][CW];
            if constexpr (LPV <= 4)
            {
                // All U x CW x h row loads of this step are issued before any is used: every lane loads without a
                // branch (a branch around the loads makes the compiler wait for them at its end), lanes without a
                // value or a word read word 0 of the matrix (always in bounds) and drop it, and a scheduling fence
                // keeps the ANDs after the last load.
                uint64_t ld[U][CW][KMAX];
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    uint32_t const idx = s + static_cast<uint32_t>(u * VPS + sub);
                    bool const act = idx < nv;
                    uint64_t r[KMAX];
#pragma unroll
                    for (int k = 0; k < KMAX; ++k)
                    {
                        r[k] = 0ULL;
                        if (HF > 0 || static_cast<uint32_t>(k) < hf)
                        {
                            if constexpr (LPV == 1 && U == 1)
                                r[k] = rows[k]; // each lane's own value (idx == lane)
                            else if constexpr (LPV == 64)
                                r[k] = readlane64(rows[k], static_cast<int>(min(idx, 63u)));
                            else
                                r[k] = shfl64(rows[k], static_cast<int>(min(idx, 63u)));
                        }
                    }
#pragma unroll
                    for (int c = 0; c < CW; ++c)
                    {
                        uint32_t const wt = wl + static_cast<uint32_t>(c * LPV); // word within tile
                        bool const live = act && wt < tw;
                        uint64_t const w = live ? w0 + wt : 0ULL;
#pragma unroll
                        for (int k = 0; k < KMAX; ++k)
                            ld[u][c][k] = (HF > 0 || static_cast<uint32_t>(k) < hf)
                                              ? load_row_word<NT>(words + (live ? r[k] : 0ULL) + w)
                                              : ~0ULL;
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    bool const act = s + static_cast<uint32_t>(u * VPS + sub) < nv;
#pragma unroll
                    for (int c = 0; c < CW; ++c)
                    {
                        uint32_t const wt = wl + static_cast<uint32_t>(c * LPV);
                        bool const live = act && wt < tw;
                        uint64_t acc = ld[u][c][0];
#pragma unroll
                        for (int k = 1; k < KMAX; ++k)
                            acc &= ld[u][c][k];
                        acc &= tail_mask(g, w0 + wt);
                        x[u][c] = live ? acc : 0ULL;
                    }
                }
            }
            else
            {
#pragma unroll
                for (int u = 0; u < U; ++u)
                {
                    uint32_t const idx = s + static_cast<uint32_t>(u * VPS + sub);
                    bool const act = idx < nv;
                    uint64_t r[KMAX];
#pragma unroll
                    for (int k = 0; k < KMAX; ++k)
                    {
                        r[k] = 0ULL;
                        if (HF > 0 || static_cast<uint32_t>(k) < hf)
                        {
                            if constexpr (LPV == 1 && U == 1)
                                r[k] = rows[k]; // each lane's own value (idx == lane)
                            else if constexpr (LPV == 64)
                                r[k] = readlane64(rows[k], static_cast<int>(min(idx, 63u)));
                            else
                                r[k] = shfl64(rows[k], static_cast<int>(min(idx, 63u)));
                        }
                    }
#pragma unroll
                    for (int c = 0; c < CW; ++c)
                    {
                        uint32_t const wt = wl + static_cast<uint32_t>(c * LPV); // word within tile
                        uint64_t acc = 0ULL;
                        if (act && wt < tw)
                        {
                            uint64_t const w = w0 + wt;
                            acc = load_row_word<NT>(words + r[0] + w);
#pragma unroll
                            for (int k = 1; k < KMAX; ++k)
                                if (HF > 0 || static_cast<uint32_t>(k) < hf)
                                    acc &= load_row_word<NT>(words + r[k] + w);
                            acc &= tail_mask(g, static_cast<uint32_t>(w));
                        }
                        x[u][c] = acc;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
            {
#pragma unroll
                for (int c = 0; c < CW; ++c)
                {
                    if constexpr (SLICED)
                    {
                        uint64_t carry = x[u][c]; // ripple-carry add of a 0/1 per bin
#pragma unroll
                        for (int p = 0; p < PLANES; ++p)
                        {
                            uint64_t const t = plane[p][c] & carry;
                            plane[p][c] ^= carry;
                            carry = t;
                        }
                    }
                    else
                    {
                        uint64_t bits = x[u][c];
                        uint32_t const cbase = (wl + static_cast<uint32_t>(c * LPV)) * 64u;
                        while (bits)
                        {
                            uint32_t const b = static_cast<uint32_t>(__builtin_ctzll(bits));
                            atomicAdd(&lds[cbase + b], 1u);
                            bits &= bits - 1ULL;
                        }
                    }
                }
            }
        }
        if constexpr (SLICED)
        {
#pragma unroll
            for (int c = 0; c < CW; ++c)
            {
                uint64_t bits = 0ULL;
#pragma unroll
                for (int p = 0; p < PLANES; ++p)
                    bits |= plane[p][c];
                uint32_t const cbase = (wl + static_cast<uint32_t>(c * LPV)) * 64u;
                while (bits)
                {
                    uint32_t const b = static_cast<uint32_t>(__builtin_ctzll(bits));
                    uint32_t n = 0;
#pragma unroll
                    for (int p = 0; p < PLANES; ++p)
                        n |= static_cast<uint32_t>((plane[p][c] >> b) & 1ULL) << p;
                    atomicAdd(&lds[cbase + b], n);
                    bits &= bits - 1ULL;
                }
            }
        }
    }
}

// Runtime dispatch of count_into_lds on log2(lanes per value): 0..6 give
// LPV = 1..64 with one word per lane (bin_words <= 64); 7 gives 64 lanes with
// two words each (bin_words <= 128, or any width over tiles of 128 words with
// TILED).  Used by kernels that visit IBFs of different widths (the HIBF
// descent).  `lds` may also be device memory.
template <int U, bool NT = false, int HF = 0, bool SLICE = true, bool TILED = false>
__device__ __forceinline__ void count_into_lds_dispatch(uint32_t lpv_log,
                                                        IbfView const & g,
                                                        uint64_t const * __restrict__ vals,
                                                        uint32_t L,
                                                        uint32_t * lds,
                                                        int wave,
                                                        int nwaves,
                                                        int lane)
{
    switch (lpv_log)
    {
    case 0: count_into_lds<1, 1, U, NT, HF, SLICE>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane); break;
    case 1: count_into_lds<2, 1, U, NT, HF, SLICE>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane); break;
    case 2: count_into_lds<4, 1, U, NT, HF, SLICE>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane); break;
    case 3: count_into_lds<8, 1, U, NT, HF, SLICE>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane); break;
    case 4: count_into_lds<16, 1, U, NT, HF, SLICE>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane); break;
    case 5: count_into_lds<32, 1, U, NT, HF, SLICE>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane); break;
    case 6: count_into_lds<64, 1, U, NT, HF, SLICE>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane); break;
    default:
        // 64 lanes with two words each cover 128 words; with TILED, wider rows are counted tile by tile (only
        // kernels whose state is in device memory meet IBFs that wide; the loop costs the others registers).
        if constexpr (TILED)
        {
            for (uint32_t w0 = 0; w0 < g.bin_words; w0 += 128u)
                count_into_lds<64, 2, U, NT, HF, SLICE>(g, vals, L, w0, min(128u, g.bin_words - w0), lds + w0 * 64u,
                                                        wave, nwaves, lane);
        }
        else
        {
            count_into_lds<64, 2, U, NT, HF, SLICE>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane);
        }
        break;
    }
}

// Dispatch for HIBFs whose IBFs all have rows of at most 4 words (256
// technical bins, the tmax = 256 layout): only the 1-, 2- and 4-word shapes are
// compiled, which keeps the kernel within 64 VGPRs.
template <int U, bool NT = false, int HF = 0>
__device__ __forceinline__ void count_into_lds_dispatch_narrow(uint32_t lpv_log,
                                                               IbfView const & g,
                                                               uint64_t const * __restrict__ vals,
                                                               uint32_t L,
                                                               uint32_t * lds,
                                                               int wave,
                                                               int nwaves,
                                                               int lane)
{
    switch (lpv_log)
    {
    case 0: count_into_lds<1, 1, U, NT, HF>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane); break;
    case 1: count_into_lds<2, 1, U, NT, HF>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane); break;
    default: count_into_lds<4, 1, U, NT, HF>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane); break;
    }
}

// Counting for HIBFs whose IBFs all have rows of at most 4 words (256
// technical bins, the tmax = 256 layout): each lane takes one value and reads
// its whole row (4 words; fewer are masked), so a wave has 64 values x h rows
// in flight at once, one memory round trip per 64 values.
template <bool NT = false, int HF = 0>
__device__ __forceinline__ void count_into_lds_row_per_lane(IbfView const & g,
                                                            uint64_t const * __restrict__ vals,
                                                            uint32_t L,
                                                            uint32_t * lds,
                                                            int wave,
                                                            int nwaves,
                                                            int lane)
{
    count_into_lds<1, 4, 1, NT, HF>(g, vals, L, 0, g.bin_words, lds, wave, nwaves, lane);
}

} // namespace hibf_amd
