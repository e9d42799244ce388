// HIBF descent kernels for gfx950 (membership_for and bulk_count), plus the
// small helpers that turn per-query hit lists into CSR output.
#include "descent_common.hpp"
#include "device_common.hpp"
#include "kernels.hpp"

namespace hibf_amd
{

namespace
{

// Shared state of one block's depth-first descent.
struct DescentState
{
    uint32_t nev[kHibfMaxDepth];    // events recorded at each level
    uint32_t pos[kHibfMaxDepth];    // next event of each level above the current one
    uint32_t ibf_at[kHibfMaxDepth]; // IBF visited at each level
    uint64_t start;                 // output slot of the current query
};

// One block per query (grid-stride).  Depth-first descent from IBF 0 with an
// explicit stack of per-level event lists in LDS, which reproduces the
// reference's recursion order (hierarchical_interleaved_bloom_filter.hpp:321-355
// and :462-496) without recursion.  HF > 0: every IBF has HF hash functions.
template <bool COUNTING, bool NT, int HF, int U = 2, int NARROW = 0, bool SLICE = true, bool GSTATE = false>
__device__ __forceinline__ void descent_body(HibfLaunch const & h,
                                             uint64_t const * __restrict__ values,
                                             uint64_t const * __restrict__ offsets,
                                             uint64_t nq,
                                             uint16_t const * __restrict__ thr16,
                                             uint64_t const * __restrict__ thr64,
                                             uint32_t counter_bytes,
                                             void * __restrict__ counts,
                                             uint32_t * __restrict__ scratch,
                                             uint64_t * __restrict__ ids,
                                             uint64_t ids_cap,
                                             uint64_t * __restrict__ hit_start,
                                             uint64_t * __restrict__ hit_count,
                                             uint64_t const * __restrict__ qlist)
{
    extern __shared__ uint32_t smem[];
    __shared__ DescentState st;
    __shared__ uint64_t next_q;
    unsigned long long * const total = h.counters;

    int const tid = static_cast<int>(threadIdx.x);
    int const nthreads = static_cast<int>(blockDim.x);
    int const lane = tid & 63;
    int const wave = tid >> 6;
    int const nwaves = nthreads >> 6;
    // The block's part of scratch: its hit list (membership), then, with GSTATE, its counters and event lists.
    uint64_t const hit_words = COUNTING ? 0u : h.max_hits;
    uint32_t * const region =
        scratch ? scratch + static_cast<uint64_t>(blockIdx.x) * (hit_words + (GSTATE ? h.gstate_words : 0u)) : nullptr;
    uint32_t * const my_scratch = COUNTING ? nullptr : region;
    uint32_t * const cnt = GSTATE ? region + hit_words : smem; // max_bins counters
    uint32_t * const evs = cnt + h.max_bins;                   // depth x max_bins event lists

    // Work stealing: each block takes the next query from the agent's cursor.
    uint64_t visits = 0;
    uint64_t row_bytes = 0;
    uint64_t lookups = 0;
    while (true)
    {
        // With a query list (the overflow pass), entry k of qlist is the query
        // and counters[6] holds the list length.
        if (tid == 0)
        {
            if (qlist)
            {
                uint64_t const k = atomicAdd(&h.counters[7], 1ULL);
                next_q = k < h.counters[6] ? qlist[k] : ~0ULL;
            }
            else
            {
                next_q = atomicAdd(&h.counters[1], 1ULL);
            }
        }
        __syncthreads();
        uint64_t const q = uniform64(next_q);
        __syncthreads();
        if (q >= nq)
            break;
        // The query's offsets and threshold (read-only inputs, uniform addresses) come through scalar loads.
        uint64_t const beg = const_load(offsets + q);
        uint32_t const L = static_cast<uint32_t>(const_load(offsets + q + 1) - beg);
        uint64_t const * const vals = values + beg;
        uint64_t const threshold = COUNTING ? const_load(thr64 + q) : static_cast<uint64_t>(thr16[q]);
        uint32_t const cb = COUNTING ? counter_bytes : 2u;
        void * const row = COUNTING ? static_cast<void *>(static_cast<char *>(counts) + q * h.n_user_bins * cb)
                                    : nullptr;
        if constexpr (COUNTING)
        {
            zero_counter_row(row, h.n_user_bins, cb, h.nt_stores != 0u, tid, nthreads);
        }
        // Every thread walks the same events, so the walk's state (level, IBF, next event, ids found) is
        // kept in registers; only a level's resume point goes through LDS, written by thread 0 when the
        // walk descends and read by all threads when it returns.
        uint32_t depth = 0;
        uint32_t ibf = 0;
        uint32_t p = 0;
        uint32_t nev = 0;
        uint32_t hits = 0;
        bool descend = true;
        while (true)
        {
            if (descend)
            {
                // Scalar loads of the descriptor where the registers allow them (the narrow-row kernels).
                HibfIbfDesc const d = NARROW == 1 ? load_desc(h.ibfs + ibf) : h.ibfs[ibf];
                for (uint32_t i = static_cast<uint32_t>(tid); i < d.bins; i += static_cast<uint32_t>(nthreads))
                    cnt[i] = 0u;
                __syncthreads();
                IbfView const g{h.arena + d.word_offset, d.bin_size, d.bins, d.bin_words, d.hash_shift, d.hash_funs};
                if constexpr (NARROW == 2)
                    count_into_lds_row_per_lane<NT, HF>(g, vals, L, cnt, wave, nwaves, lane);
                else if constexpr (NARROW == 1)
                {
                    // IBFs too large for the caches to keep rows between queries stream their rows.
                    if (!NT && static_cast<uint64_t>(d.bin_size) * d.bin_words * 8u >= h.stream_bytes)
                        count_into_lds_dispatch_narrow<U, true, HF>(d.lpv_log, g, vals, L, cnt, wave, nwaves, lane);
                    else
                        count_into_lds_dispatch_narrow<U, NT, HF>(d.lpv_log, g, vals, L, cnt, wave, nwaves, lane);
                }
                // (Not in the 8-wave two-hash membership kernel: a second row-load form does not fit its 64 VGPRs.)
                else if (!NT && !GSTATE && (COUNTING || HF != 2)
                         && static_cast<uint64_t>(d.bin_size) * d.bin_words * 8u >= h.stream_bytes)
                    count_into_lds_dispatch<U, true, HF, SLICE, GSTATE>(d.lpv_log, g, vals, L, cnt, wave, nwaves, lane);
                else
                    count_into_lds_dispatch<U, NT, HF, SLICE, GSTATE>(d.lpv_log, g, vals, L, cnt, wave, nwaves, lane);
                ++visits;
                lookups += L;
                row_bytes += static_cast<uint64_t>(L) * d.hash_funs * d.bin_words * 8u;
                __syncthreads();
                if (wave == 0)
                {
                    uint32_t const n =
                        scan_ibf<COUNTING>(d, h.tb_info, cnt, threshold, cb, evs + depth * h.max_bins, row, lane);
                    if (lane == 0)
                        st.nev[depth] = n;
                }
                __syncthreads();
                nev = uniform32(st.nev[depth]);
                p = 0;
                descend = false;
            }
            if (p < nev)
            {
                // Next event of the current level, in technical-bin order.
                uint32_t const e = uniform32(evs[depth * h.max_bins + p]);
                ++p;
                if (e & kEvMerged)
                {
                    if (tid == 0)
                    {
                        st.pos[depth] = p;
                        st.nev[depth] = nev;
                        st.ibf_at[depth] = ibf;
                    }
                    ++depth;
                    ibf = e & ~kEvMerged;
                    descend = true;
                }
                else if constexpr (!COUNTING)
                {
                    // max_hits bounds the run ends of one descent; the check only guards the list.
                    if (tid == 0)
                    {
                        if (hits < h.max_hits)
                            my_scratch[hits] = e;
                        else
                            h.counters[5] = 1ULL;
                    }
                    ++hits;
                }
                continue;
            }
            // Level exhausted: return to the parent.  The barrier keeps thread 0 from overwriting the parent's
            // resume point (at its next merged bin) before every thread has read it.
            if (depth == 0)
                break;
            --depth;
            ibf = uniform32(st.ibf_at[depth]);
            p = uniform32(st.pos[depth]);
            nev = uniform32(st.nev[depth]);
            __syncthreads();
        }

        if constexpr (!COUNTING)
        {
            hits = min(hits, static_cast<uint32_t>(h.max_hits));
            __syncthreads();
            uint32_t const cap = h.per_query_cap;
            if (tid == 0)
            {
                uint64_t const found = atomicAdd(total, static_cast<unsigned long long>(hits));
                uint64_t const s = cap ? q * cap : found;
                uint32_t const kept = cap ? min(hits, cap) : hits;
                st.start = s;
                hit_start[q] = s;
                hit_count[q] = hits;
                if (s + kept > ids_cap)
                    h.counters[5] = 1ULL;
            }
            __syncthreads();
            uint32_t const n = cap ? min(hits, cap) : hits;
            uint64_t const s = st.start;
            for (uint32_t i = static_cast<uint32_t>(tid); i < n; i += static_cast<uint32_t>(nthreads))
                if (s + i < ids_cap)
                    ids[s + i] = my_scratch[i];
        }
        __syncthreads();
    }
    if (tid == 0 && visits)
    {
        atomicAdd(&h.counters[2], static_cast<unsigned long long>(visits));
        atomicAdd(&h.counters[3], static_cast<unsigned long long>(row_bytes));
        atomicAdd(&h.counters[4], static_cast<unsigned long long>(lookups));
    }
}

#define HIBF_DESCENT_PARAMS                                                                                          \
    HibfLaunch h, uint64_t const * __restrict__ values, uint64_t const * __restrict__ offsets, uint64_t nq,            \
        uint16_t const * __restrict__ thr16, uint64_t const * __restrict__ thr64, uint32_t counter_bytes,               \
        void * __restrict__ counts, uint32_t * __restrict__ scratch, uint64_t * __restrict__ ids, uint64_t ids_cap,     \
        uint64_t * __restrict__ hit_start, uint64_t * __restrict__ hit_count, uint64_t const * __restrict__ qlist
#define HIBF_DESCENT_ARGS                                                                                            \
    h, values, offsets, nq, thr16, thr64, counter_bytes, counts, scratch, ids, ids_cap, hit_start, hit_count, qlist

template <bool COUNTING, bool NT, int HF>
__global__ void __launch_bounds__(512) hibf_descent_kernel(HIBF_DESCENT_PARAMS)
{
    descent_body<COUNTING, NT, HF>(HIBF_DESCENT_ARGS);
}

// HIBFs whose counters and event lists do not fit the LDS budget (IBFs of
// more than a few thousand technical bins, deep trees): the same descent with
// that state in the block's region of device memory.
template <bool COUNTING>
__global__ void __launch_bounds__(512) hibf_descent_gstate_kernel(HIBF_DESCENT_PARAMS)
{
    descent_body<COUNTING, false, 0, 2, 0, true, true>(HIBF_DESCENT_ARGS);
}

// Membership with two hash functions (the HIBF default) fits 64 VGPRs without
// spilling, so it runs at 8 waves per SIMD: more row fetches in flight.
// NT: non-temporal row loads.
template <bool NT = false>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8, 8)))
hibf_descent_membership_h2_kernel(HIBF_DESCENT_PARAMS)
{
    descent_body<false, NT, 2, 2, 0, false>(HIBF_DESCENT_ARGS);
}

// Same, for HIBFs whose rows are at most 4 words: one value and its whole row per lane.
template <bool NT = false>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8, 8)))
hibf_descent_membership_h2_row_per_lane_kernel(HIBF_DESCENT_PARAMS)
{
    descent_body<false, NT, 2, 1, 2>(HIBF_DESCENT_ARGS);
}

// Membership with two hash functions over rows of at most 4 words (the
// tmax = 256 layout): only the 1-, 2- and 4-word counting shapes are compiled.
// SLICE: bit-sliced per-lane counts with one LDS atomic per bin and chunk.
// U: value groups whose rows a wave loads at once (2: 32 values of 4-word rows,
// 4: 64 values, a whole chunk per memory round trip).  NT: non-temporal row loads.
template <bool SLICE, int U = 2, bool NT = false>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8, 8)))
hibf_descent_membership_h2_w4_kernel(HIBF_DESCENT_PARAMS)
{
    descent_body<false, NT, 2, U, 1, SLICE>(HIBF_DESCENT_ARGS);
}

using DescentKernel = void (*)(HibfLaunch, uint64_t const *, uint64_t const *, uint64_t, uint16_t const *,
                               uint64_t const *, uint32_t, void *, uint32_t *, uint64_t *, uint64_t, uint64_t *,
                               uint64_t *, uint64_t const *);

// The membership kernel of an HIBF under the current launch-shape knobs.
template <bool NT>
DescentKernel membership_kernel(HibfLaunch const & h)
{
    bool const h2 = h.uniform_hash_funs == 2u;
    bool const w4 = h2 && h.max_bin_words <= 4u;
    CountTuning const & t = g_count_tuning;
    if (h.gstate_words)
        return hibf_descent_gstate_kernel<false>;
    if (w4 && t.row_per_lane)
        return hibf_descent_membership_h2_row_per_lane_kernel<NT>;
    if (w4 && t.sliced_counts)
        return t.hibf_unroll == 4u ? hibf_descent_membership_h2_w4_kernel<true, 4, NT>
                                   : hibf_descent_membership_h2_w4_kernel<true, 2, NT>;
    if (w4)
        return hibf_descent_membership_h2_w4_kernel<false, 2, NT>;
    if (h2)
        return hibf_descent_membership_h2_kernel<NT>;
    return hibf_descent_kernel<false, NT, 0>;
}

#undef HIBF_DESCENT_PARAMS
#undef HIBF_DESCENT_ARGS

__global__ void reset_counters_kernel(unsigned long long * counters, uint32_t mask)
{
    if (threadIdx.x < 32u && ((mask >> threadIdx.x) & 1u))
        counters[threadIdx.x] = 0ULL;
}

__global__ void __launch_bounds__(256) gather_hits_kernel(uint64_t const * __restrict__ src,
                                                          uint64_t const * __restrict__ start,
                                                          uint64_t const * __restrict__ count,
                                                          uint64_t const * __restrict__ offsets,
                                                          uint64_t nq,
                                                          uint64_t * __restrict__ out)
{
    for (uint64_t q = blockIdx.x; q < nq; q += gridDim.x)
    {
        uint64_t const n = count[q];
        uint64_t const s = start[q];
        uint64_t const o = offsets[q];
        for (uint64_t i = threadIdx.x; i < n; i += blockDim.x)
            out[o + i] = src[s + i];
    }
}

// IBF membership_for from uint16 counts (interleaved_bloom_filter.hpp:757-761),
// one wave per query.  threshold_mask_kernel reads a query's counts once, 8
// counters per lane and 16-byte load, and writes one bit per bin (count >= t)
// into a mask of mask_words 64-bit words (bit b of the mask = bin b) plus the
// query's hit count.  mask_compact_kernel, after the scan of the hit counts,
// writes the set bins of each mask in ascending order from hit_offsets[q]:
// lane j owns mask word j of a 64-word stretch, and a wave prefix sum of the
// words' popcounts places each lane's bins.
__global__ void __launch_bounds__(256) threshold_mask_kernel(uint16_t const * __restrict__ counts,
                                                             uint64_t bins,
                                                             uint64_t nq,
                                                             uint16_t const * __restrict__ thresholds,
                                                             uint64_t * __restrict__ hit_count,
                                                             uint8_t * __restrict__ mask,
                                                             uint64_t mask_words)
{
    int const lane = static_cast<int>(threadIdx.x) & 63;
    uint64_t const waves_per_block = blockDim.x / 64u;
    uint64_t const stride = static_cast<uint64_t>(gridDim.x) * waves_per_block;
    uint64_t const mask_bytes = mask_words * 8u;
    bool const vec = (bins % 8u) == 0u; // then every query's counts start 16-byte aligned
    for (uint64_t q = blockIdx.x * waves_per_block + (threadIdx.x >> 6); q < nq; q += stride)
    {
        uint32_t const t = thresholds[q];
        uint16_t const * const c = counts + q * bins;
        uint8_t * const m = mask + q * mask_bytes;
        uint32_t hits = 0;
        for (uint64_t g = static_cast<uint64_t>(lane); g < mask_bytes; g += 64u) // g: bins 8g .. 8g + 7
        {
            uint64_t const b0 = g * 8u;
            uint32_t byte = 0;
            if (vec && b0 < bins)
            {
                uint4 const w = *reinterpret_cast<uint4 const *>(c + b0);
                uint32_t const x[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int j = 0; j < 4; ++j)
                {
                    byte |= ((x[j] & 0xFFFFu) >= t ? 1u : 0u) << (2 * j);
                    byte |= ((x[j] >> 16) >= t ? 1u : 0u) << (2 * j + 1);
                }
            }
            else
            {
                for (int j = 0; j < 8; ++j)
                    if (b0 + j < bins && c[b0 + j] >= t)
                        byte |= 1u << j;
            }
            m[g] = static_cast<uint8_t>(byte);
            hits += static_cast<uint32_t>(__popc(byte));
        }
        for (int o = 32; o > 0; o >>= 1)
            hits += __shfl_xor(hits, o);
        if (lane == 0)
            hit_count[q] = hits;
    }
}

__global__ void __launch_bounds__(256) mask_compact_kernel(uint64_t const * __restrict__ mask,
                                                           uint64_t mask_words,
                                                           uint64_t nq,
                                                           uint64_t const * __restrict__ hit_offsets,
                                                           uint64_t * __restrict__ ids,
                                                           uint64_t ids_cap)
{
    int const lane = static_cast<int>(threadIdx.x) & 63;
    uint64_t const waves_per_block = blockDim.x / 64u;
    uint64_t const stride = static_cast<uint64_t>(gridDim.x) * waves_per_block;
    for (uint64_t q = blockIdx.x * waves_per_block + (threadIdx.x >> 6); q < nq; q += stride)
    {
        uint64_t o = hit_offsets[q];
        uint64_t const * const m = mask + q * mask_words;
        for (uint64_t w0 = 0; w0 < mask_words; w0 += 64u)
        {
            uint64_t const wi = w0 + static_cast<uint64_t>(lane);
            uint64_t bits = wi < mask_words ? m[wi] : 0ULL;
            uint32_t const n = static_cast<uint32_t>(__popcll(bits));
            uint32_t incl = n;
            for (int d = 1; d < 64; d <<= 1)
            {
                uint32_t const y = __shfl_up(incl, d);
                if (lane >= d)
                    incl += y;
            }
            uint64_t p = o + (incl - n);
            while (bits)
            {
                if (p < ids_cap)
                    ids[p] = wi * 64u + static_cast<uint64_t>(__builtin_ctzll(bits));
                ++p;
                bits &= bits - 1ULL;
            }
            o += __shfl(incl, 63);
        }
    }
}

__global__ void __launch_bounds__(256) finalize_offsets_kernel(uint64_t const * __restrict__ incl,
                                                               uint64_t n,
                                                               uint64_t base,
                                                               uint64_t const * d_base,
                                                               bool write_first,
                                                               uint64_t * out)
{
    uint64_t const i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    uint64_t const b = base + (d_base ? *d_base : 0u);
    if (i == 0 && write_first)
        out[0] = b;
    if (i < n)
        out[1 + i] = b + incl[i];
}

} // namespace

hipError_t launch_finalize_offsets(uint64_t const * incl, uint64_t n, uint64_t base, uint64_t const * d_base,
                                   bool write_first, uint64_t * out, hipStream_t stream)
{
    unsigned const blocks = static_cast<unsigned>((n + 255u) / 256u + 1u);
    hipLaunchKernelGGL(finalize_offsets_kernel, dim3(blocks), dim3(256), 0, stream, incl, n, base, d_base, write_first,
                       out);
    return hipGetLastError();
}

hipError_t launch_reset_counters(unsigned long long * counters, uint32_t mask, hipStream_t stream)
{
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, stream, counters, mask);
    return hipGetLastError();
}

hipError_t launch_hibf_membership(HibfLaunch const & h,
                                  uint64_t const * values,
                                  uint64_t const * offsets,
                                  uint64_t nq,
                                  uint16_t const * thresholds,
                                  uint32_t * scratch,
                                  unsigned blocks,
                                  uint64_t * ids_tmp,
                                  uint64_t ids_cap,
                                  uint64_t * hit_start,
                                  uint64_t * hit_count,
                                  uint64_t const * qlist,
                                  hipStream_t stream)
{
    if (nq == 0)
        return hipSuccess;
    size_t const lds = h.gstate_words ? 0u : static_cast<size_t>(h.max_bins) * (1u + h.depth) * sizeof(uint32_t);
    DescentKernel const kernel = g_count_tuning.nt_loads ? membership_kernel<true>(h) : membership_kernel<false>(h);
    unsigned const threads = 64u * g_count_tuning.hibf_block_waves;
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), lds, stream, h, values, offsets, nq, thresholds, nullptr,
                       2u, nullptr, scratch, ids_tmp, ids_cap, hit_start, hit_count, qlist);
    return hipGetLastError();
}

hipError_t launch_hibf_count(HibfLaunch const & h,
                             uint64_t const * values,
                             uint64_t const * offsets,
                             uint64_t nq,
                             uint64_t const * thresholds,
                             uint32_t counter_bytes,
                             void * counts,
                             uint32_t * scratch,
                             unsigned blocks,
                             hipStream_t stream)
{
    if (nq == 0)
        return hipSuccess;
    size_t const lds = h.gstate_words ? 0u : static_cast<size_t>(h.max_bins) * (1u + h.depth) * sizeof(uint32_t);
    auto const kernel = h.gstate_words               ? hibf_descent_gstate_kernel<true>
                        : h.uniform_hash_funs == 2u ? hibf_descent_kernel<true, false, 2>
                                                    : hibf_descent_kernel<true, false, 0>;
    hipLaunchKernelGGL(kernel, dim3(blocks), dim3(256), lds, stream, h, values, offsets, nq, nullptr, thresholds,
                       counter_bytes, counts, scratch, nullptr, 0ULL, nullptr, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t launch_gather_hits(uint64_t const * src,
                              uint64_t const * start,
                              uint64_t const * count,
                              uint64_t const * offsets,
                              uint64_t nq,
                              uint64_t * out,
                              hipStream_t stream)
{
    if (nq == 0)
        return hipSuccess;
    unsigned const blocks = static_cast<unsigned>(nq < 65536u ? nq : 65536u);
    hipLaunchKernelGGL(gather_hits_kernel, dim3(blocks), dim3(256), 0, stream, src, start, count, offsets, nq, out);
    return hipGetLastError();
}

hipError_t launch_threshold_mask(uint16_t const * counts,
                                 uint64_t bins,
                                 uint64_t nq,
                                 uint16_t const * thresholds,
                                 uint64_t * hit_count,
                                 uint64_t * mask,
                                 uint64_t mask_words,
                                 hipStream_t stream)
{
    if (nq == 0)
        return hipSuccess;
    unsigned const blocks = static_cast<unsigned>(std::min<uint64_t>((nq + 3u) / 4u, 1u << 16));
    hipLaunchKernelGGL(threshold_mask_kernel, dim3(blocks), dim3(256), 0, stream, counts, bins, nq, thresholds,
                       hit_count, reinterpret_cast<uint8_t *>(mask), mask_words);
    return hipGetLastError();
}

hipError_t launch_mask_compact(uint64_t const * mask,
                               uint64_t mask_words,
                               uint64_t nq,
                               uint64_t const * hit_offsets,
                               uint64_t * ids,
                               uint64_t ids_cap,
                               hipStream_t stream)
{
    if (nq == 0)
        return hipSuccess;
    unsigned const blocks = static_cast<unsigned>(std::min<uint64_t>((nq + 3u) / 4u, 1u << 16));
    hipLaunchKernelGGL(mask_compact_kernel, dim3(blocks), dim3(256), 0, stream, mask, mask_words, nq, hit_offsets, ids,
                       ids_cap);
    return hipGetLastError();
}

} // namespace hibf_amd
