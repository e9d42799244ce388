/*
 * Oracle for the HIBF query path: a plain-C restatement of the reference
 * algorithm.  TEST INFRASTRUCTURE ONLY.  It is linked by tests/, by the smoke
 * check and by bench.py's CPU-baseline leg, always as the checker and never
 * as the thing that is measured or shipped.  The product (hibf_amd/) never
 * links it.
 *
 * Each function cites the reference lines it restates (paths relative to the
 * seqan/hibf 1.0.0-rc.1 source tree).
 *
 * Pinning (what checks each function beyond retyping its formula):
 *   orc_hash_and_fit       Python big-integer restatement + committed vectors
 *                          + closed-form special cases (tests/test_oracle_hash.py).
 *   orc_ibf_init/emplace   snippet goldens (counting_agent.out, ...), the
 *                          insert_iterator unit test counts, brute force
 *                          "h bits set in column b only".
 *   orc_ibf_bulk_contains  containment_agent_bulk_contains.out, invariants of
 *                          interleaved_bloom_filter_test.cpp (no false negatives).
 *   orc_ibf_bulk_count     counting_agent.out (uint16 and uint8), counting_vector.out,
 *                          counting_vector_test.cpp wrap-around cases.
 *   orc_ibf_membership_for membership_agent_membership_for.out ([3,5,7]).
 *   orc_hibf_membership_for / orc_hibf_bulk_count
 *                          hierarchical_interleaved_bloom_filter.out, readme.out,
 *                          the reference HIBF unit-test expectations, and a
 *                          brute-force tree check (a user bin is reported iff its
 *                          run sum and every ancestor merged-bin sum reach t).
 *   orc_hll_*              hyperloglog_test.cpp expectations and exact counts
 *                          (see the HyperLogLog section).
 *   orc_mt19937_64         the C++ standard's 10,000th output for seed 5489.
 *   orc_synth_value        test-input generator, not reference code: its
 *                          Python twin (hibf_amd/synth.py) and the device
 *                          stream must agree (tests/test_oracle.py).
 *   orc_ibf_emplace_*      parallel forms of orc_ibf_emplace: equal to it on
 *                          the same pairs (tests/test_oracle.py).
 *   orc_fs_*               golden iteration orders written by the reference's
 *                          own robin_hood set (tests/golden/robin_hood_orders.json)
 *                          and, where the reference tree is present, that set
 *                          run on random scripts (tests/test_split_order.py).
 *   orc_minimiser_hash     seqan3's minimiser_hash (a dependency of the
 *                          reference's tutorial, absent here): an independent
 *                          chain form of the rule, closed forms and invariants
 *                          (tests/test_oracle.py); parity with seqan3's own
 *                          output is unpinned.
 */
#include <math.h>
#include <stdio.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* interleaved_bloom_filter.hpp:172-176 */
static const uint64_t ORC_HASH_SEEDS[5] = {13572355802537770549ULL, 13043817825332782213ULL,
                                           10650232656628343401ULL, 16499269484942379435ULL,
                                           4893150838803335377ULL};

/* hierarchical_interleaved_bloom_filter.hpp:32-40 */
#define ORC_BIN_MERGED 0xFFFFFFFFFFFFFFFFULL
#define ORC_BIN_DELETED 0xFFFFFFFFFFFFFFFEULL

/* Geometry of one IBF: interleaved_bloom_filter.hpp:160-170. */
typedef struct orc_ibf
{
    uint64_t bins;
    uint64_t technical_bins;
    uint64_t bin_size;
    uint64_t hash_shift;
    uint64_t bin_words;
    uint64_t hash_funs;
} orc_ibf;

uint64_t orc_hash_seed(uint32_t i) { return i < 5 ? ORC_HASH_SEEDS[i] : 0; }

/* misc/divide_and_ceil.hpp:20-26 */
static uint64_t orc_divide_and_ceil(uint64_t a, uint64_t b) { return (a + (b - 1u)) / b; }

static uint64_t orc_countl_zero(uint64_t x)
{
    uint64_t n = 0;
    if (x == 0)
        return 64;
    while (!(x & 0x8000000000000000ULL))
    {
        x <<= 1;
        ++n;
    }
    return n;
}

/* src/interleaved_bloom_filter.cpp:32-53.  Returns 0, or 1/2/3 for the three
 * logic_error cases in the order the reference checks them. */
int orc_ibf_init(orc_ibf * g, uint64_t bins, uint64_t bin_size, uint64_t hash_funs)
{
    if (bins == 0)
        return 1;
    if (hash_funs == 0 || hash_funs > 5)
        return 2;
    if (bin_size == 0)
        return 3;
    g->bins = bins;
    g->bin_size = bin_size;
    g->hash_funs = hash_funs;
    g->hash_shift = orc_countl_zero(bin_size);
    g->bin_words = orc_divide_and_ceil(bins, 64u);
    g->technical_bins = g->bin_words * 64u;
    return 0;
}

uint64_t orc_ibf_word_count(orc_ibf const * g) { return g->technical_bins * g->bin_size / 64u; }

/* interleaved_bloom_filter.hpp:185-195: seeded multiply, xor-shift, golden-ratio
 * multiply, fastrange (high half of the 128-bit product), times technical_bins. */
uint64_t orc_hash_and_fit(uint64_t h, uint64_t seed, uint64_t hash_shift, uint64_t bin_size, uint64_t technical_bins)
{
    h *= seed;
    h ^= h >> hash_shift;
    h *= 11400714819323198485ULL;
    h = (uint64_t)(((unsigned __int128)h * (unsigned __int128)bin_size) >> 64);
    h *= technical_bins;
    return h;
}

/* src/interleaved_bloom_filter.cpp:119-140 with track_occupancy == false:
 * set bit hash_and_fit(v, seed_i) + bin for every i < h.  Bits are LSB-first
 * inside 64-bit words (bit_vector.hpp:1045-1054). */
void orc_ibf_emplace(uint64_t * words, orc_ibf const * g, uint64_t const * values, uint64_t const * bins, uint64_t n)
{
    for (uint64_t k = 0; k < n; ++k)
    {
        for (uint64_t i = 0; i < g->hash_funs; ++i)
        {
            uint64_t idx = orc_hash_and_fit(values[k], ORC_HASH_SEEDS[i], g->hash_shift, g->bin_size, g->technical_bins)
                         + bins[k];
            words[idx >> 6] |= 1ULL << (idx & 63u);
        }
    }
}

/* src/interleaved_bloom_filter.cpp:119-140 with track_occupancy == true, for
 * the pairs k = 0, 1, ... in order: exists starts true and is and-ed with each
 * hash's bit read just before that bit is set (so a value whose two hashes hit
 * one unset bit still finds it unset at the first read); occupancy[bin] grows
 * by one iff !exists.  Pairs with bins[k] >= g->bins are skipped (the
 * reference asserts bin < bins). */
void orc_ibf_emplace_tracked(uint64_t * words, orc_ibf const * g, uint64_t const * values, uint64_t const * bins,
                             uint64_t n, uint64_t * occupancy)
{
    for (uint64_t k = 0; k < n; ++k)
    {
        if (bins[k] >= g->bins)
            continue;
        int exists = 1;
        for (uint64_t i = 0; i < g->hash_funs; ++i)
        {
            uint64_t idx = orc_hash_and_fit(values[k], ORC_HASH_SEEDS[i], g->hash_shift, g->bin_size, g->technical_bins)
                         + bins[k];
            uint64_t const bit = 1ULL << (idx & 63u);
            exists &= (words[idx >> 6] & bit) != 0;
            words[idx >> 6] |= bit;
        }
        if (!exists)
            ++occupancy[bins[k]];
    }
}

/* Test-input generator (not reference code): value i of the seeded synthetic
 * set `bin`, the splitmix/murmur 64-bit finaliser of
 * (bin << 36 | i) + seed * 0x9E3779B97F4A7C15.  The tests draw user bins and
 * technical bins from this stream; the device builds use the same stream. */
uint64_t orc_synth_value(uint64_t seed, uint64_t bin, uint64_t i)
{
    uint64_t k = ((bin << 36) | i) + seed * 0x9E3779B97F4A7C15ULL;
    k ^= k >> 33;
    k *= 0xFF51AFD7ED558CCDULL;
    k ^= k >> 33;
    k *= 0xC4CEB9FE1A85EC53ULL;
    k ^= k >> 33;
    return k;
}

/* The reference's config constructor fills an IBF with input_fn(i) for every
 * bin in parallel (src/interleaved_bloom_filter.cpp:104-117); here bin
 * first_bin + b receives orc_synth_value(seed, first_bin + b, j) for
 * j < per_bin.  Threads own whole 64-bin words (the reference's rule for
 * concurrent emplace, interleaved_bloom_filter.hpp:142-144), so the result is
 * the sequential emplace's. */
void orc_ibf_emplace_synthetic(uint64_t * words, orc_ibf const * g, uint64_t seed, uint64_t first_bin,
                               uint64_t n_bins, uint64_t per_bin, int threads)
{
    uint64_t const w0 = first_bin / 64u;
    uint64_t const w1 = n_bins ? (first_bin + n_bins - 1u) / 64u + 1u : w0;
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
#endif
    for (int64_t w = (int64_t)w0; w < (int64_t)w1; ++w)
    {
        uint64_t const lo = (uint64_t)w * 64u < first_bin ? first_bin : (uint64_t)w * 64u;
        uint64_t const hi = ((uint64_t)w + 1u) * 64u < first_bin + n_bins ? ((uint64_t)w + 1u) * 64u : first_bin + n_bins;
        for (uint64_t b = lo; b < hi; ++b)
            for (uint64_t j = 0; j < per_bin; ++j)
            {
                uint64_t const v = orc_synth_value(seed, b, j);
                orc_ibf_emplace(words, g, &v, &b, 1u);
            }
    }
}

/* Emplace of generated ranges: item k puts orc_synth_value(seed, set[k],
 * begin[k] + j), j < count[k], into bin bins[k].  Items run on `threads`
 * threads with atomic OR on the words, so the result is the sequential
 * emplace's whatever the order. */
void orc_ibf_emplace_synthetic_items(uint64_t * words, orc_ibf const * g, uint64_t seed, uint64_t const * set,
                                     uint64_t const * begin, uint64_t const * count, uint64_t const * bins,
                                     uint64_t n_items, int threads)
{
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1)
#endif
    for (int64_t k = 0; k < (int64_t)n_items; ++k)
        for (uint64_t j = 0; j < count[k]; ++j)
        {
            uint64_t const v = orc_synth_value(seed, set[k], begin[k] + j);
            for (uint64_t i = 0; i < g->hash_funs; ++i)
            {
                uint64_t const idx =
                    orc_hash_and_fit(v, ORC_HASH_SEEDS[i], g->hash_shift, g->bin_size, g->technical_bins) + bins[k];
                __atomic_fetch_or(&words[idx >> 6], 1ULL << (idx & 63u), __ATOMIC_RELAXED);
            }
        }
}

/* orc_ibf_emplace over (value, bin) pairs on `threads` threads with atomic OR. */
void orc_ibf_emplace_parallel(uint64_t * words, orc_ibf const * g, uint64_t const * values, uint64_t const * bins,
                              uint64_t n, int threads)
{
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
#endif
    for (int64_t k = 0; k < (int64_t)n; ++k)
        for (uint64_t i = 0; i < g->hash_funs; ++i)
        {
            uint64_t const idx =
                orc_hash_and_fit(values[k], ORC_HASH_SEEDS[i], g->hash_shift, g->bin_size, g->technical_bins) + bins[k];
            __atomic_fetch_or(&words[idx >> 6], 1ULL << (idx & 63u), __ATOMIC_RELAXED);
        }
}

/* src/interleaved_bloom_filter.cpp:142-148: clear(bin) zeroes the bin's column. */
void orc_ibf_clear(uint64_t * words, orc_ibf const * g, uint64_t bin)
{
    for (uint64_t idx = bin, i = 0; i < g->bin_size; idx += g->technical_bins, ++i)
        words[idx >> 6] &= ~(1ULL << (idx & 63u));
}

/* src/interleaved_bloom_filter.cpp:150-201: increase_bin_number_to.  When the
 * words per row stay the same only `bins` changes (try_increase_bin_number_to).
 * Otherwise every row moves, last row first and last word first, from its
 * old stride to the new one in place; each moved word's old place is zeroed,
 * so the new columns end up zero.  `words` must hold the new matrix size with
 * the old matrix in front and zeros behind it.  Returns 1 when new_bins is
 * below bins (the reference throws std::invalid_argument). */
int orc_ibf_increase_bin_number_to(uint64_t * words, orc_ibf * g, uint64_t new_bins)
{
    if (new_bins < g->bins)
        return 1;
    uint64_t const old_bw = g->bin_words;
    uint64_t const new_bw = (new_bins + 63u) / 64u;
    if (new_bw > old_bw)
    {
        for (uint64_t r = g->bin_size; r-- > 1u;)
            for (uint64_t w = old_bw; w-- > 0u;)
            {
                words[r * new_bw + w] = words[r * old_bw + w];
                words[r * old_bw + w] = 0u;
            }
        g->bin_words = new_bw;
        g->technical_bins = new_bw * 64u;
    }
    g->bins = new_bins;
    return 0;
}

/* src/interleaved_bloom_filter.cpp:204-291: copy row 0, AND rows 1..h-1.
 * out receives bin_words words. */
void orc_ibf_bulk_contains(uint64_t const * words, orc_ibf const * g, uint64_t value, uint64_t * out)
{
    uint64_t idx[5];
    for (uint64_t i = 0; i < g->hash_funs; ++i)
        idx[i] = orc_hash_and_fit(value, ORC_HASH_SEEDS[i], g->hash_shift, g->bin_size, g->technical_bins) / 64u;
    memcpy(out, words + idx[0], sizeof(uint64_t) * g->bin_words);
    for (uint64_t i = 1; i < g->hash_funs; ++i)
        for (uint64_t j = 0; j < g->bin_words; ++j)
            out[j] &= words[idx[i] + j];
}

/* Store/load a counter of `width` bytes (1, 2, 4 or 8), wrapping modulo
 * 2^(8*width) like counting_vector<uintN_t>. */
static uint64_t orc_width_mask(int width) { return width >= 8 ? ~0ULL : ((1ULL << (8 * width)) - 1u); }

static uint64_t orc_load(void const * p, uint64_t i, int width)
{
    switch (width)
    {
    case 1: return ((uint8_t const *)p)[i];
    case 2: return ((uint16_t const *)p)[i];
    case 4: return ((uint32_t const *)p)[i];
    default: return ((uint64_t const *)p)[i];
    }
}

static void orc_store(void * p, uint64_t i, int width, uint64_t v)
{
    switch (width)
    {
    case 1: ((uint8_t *)p)[i] = (uint8_t)v; break;
    case 2: ((uint16_t *)p)[i] = (uint16_t)v; break;
    case 4: ((uint32_t *)p)[i] = (uint32_t)v; break;
    default: ((uint64_t *)p)[i] = v; break;
    }
}

#if defined(__x86_64__)
#include <immintrin.h>
/* counting_vector.hpp:252-278, the path a reference build takes with
 * HIBF_HAS_AVX512 (-march=native on an AVX-512 host), for uint16_t counters:
 * 32 counters per 512-bit register; the 32 bits of the bit vector that cover
 * them pick the lanes that get 1 (maskz_mov of set1(1)), added with
 * wrap-around.  The reference loads and stores whole registers of its padded
 * counter buffer; here the last chunk loads and stores only the counters below
 * nbits, which gives the same counts. */
__attribute__((target("avx512f,avx512bw"))) static void orc_counting_add_u16_avx512(uint16_t * counts,
                                                                                   uint64_t const * bits,
                                                                                   uint64_t nbits)
{
    uint32_t const * mask = (uint32_t const *)bits;
    __m512i const one = _mm512_set1_epi16(1);
    uint64_t const chunks = (nbits + 31u) / 32u;
    for (uint64_t c = 0; c < chunks; ++c)
    {
        __mmask32 const live = (c + 1u) * 32u <= nbits ? (__mmask32)0xFFFFFFFFu
                                                      : (__mmask32)((1ULL << (nbits - c * 32u)) - 1u);
        __m512i v = _mm512_maskz_loadu_epi16(live, counts + c * 32u);
        v = _mm512_add_epi16(v, _mm512_maskz_mov_epi16((__mmask32)mask[c], one));
        _mm512_mask_storeu_epi16(counts + c * 32u, live, v);
    }
}
#endif

/* 1: uint16_t counting uses the AVX-512 restatement above (set by
 * orc_set_simd_counting when the CPU has AVX-512BW). */
static int orc_simd_counting = 0;

/* Selects the counting path of uint16_t counters: on != 0 and a CPU with
 * AVX-512BW gives the AVX-512 restatement, otherwise the scalar one.  Returns
 * the path now in use (1 = AVX-512). */
int orc_set_simd_counting(int on)
{
#if defined(__x86_64__)
    __builtin_cpu_init();
    orc_simd_counting = on && __builtin_cpu_supports("avx512bw") && __builtin_cpu_supports("avx512f");
#else
    (void)on;
    orc_simd_counting = 0;
#endif
    return orc_simd_counting;
}

/* counting_vector.hpp:171-175 and the scalar impl :279-311: for each set bit
 * b of the bit vector (b < nbits), counts[b] += 1 modulo 2^(8*width). */
void orc_counting_vector_add(void * counts, int width, uint64_t const * bits, uint64_t nbits)
{
#if defined(__x86_64__)
    if (orc_simd_counting && width == 2)
    {
        orc_counting_add_u16_avx512((uint16_t *)counts, bits, nbits);
        return;
    }
#endif
    uint64_t words = orc_divide_and_ceil(nbits, 64u);
    for (uint64_t w = 0; w < words; ++w)
    {
        uint64_t x = bits[w];
        while (x)
        {
            uint64_t b = w * 64u + (uint64_t)__builtin_ctzll(x);
            x &= x - 1u;
            if (b < nbits)
                orc_store(counts, b, width, orc_load(counts, b, width) + 1u);
        }
    }
}

/* interleaved_bloom_filter.hpp:652-665: zero the counters, then add
 * bulk_contains(v) for every value.  counts holds g->bins counters. */
void orc_ibf_bulk_count(uint64_t const * words,
                        orc_ibf const * g,
                        uint64_t const * values,
                        uint64_t n,
                        int width,
                        void * counts)
{
    uint64_t row[4096];
    uint64_t * buf = g->bin_words <= 4096 ? row : (uint64_t *)malloc(8u * g->bin_words);
    memset(counts, 0, (size_t)g->bins * (size_t)width);
    for (uint64_t k = 0; k < n; ++k)
    {
        orc_ibf_bulk_contains(words, g, values[k], buf);
        orc_counting_vector_add(counts, width, buf, g->bins);
    }
    if (buf != row)
        free(buf);
}

/* Batched form over CSR queries (query q = values[offsets[q] .. offsets[q+1])),
 * one agent per thread with static scheduling as the reference documents for
 * multi-threaded callers (interleaved_bloom_filter.hpp:547-551).
 * counts is nq x g->bins counters of `width` bytes. */
void orc_ibf_bulk_count_batch(uint64_t const * words,
                              orc_ibf const * g,
                              uint64_t const * values,
                              uint64_t const * offsets,
                              uint64_t nq,
                              int width,
                              void * counts,
                              int threads)
{
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
#endif
    for (int64_t q = 0; q < (int64_t)nq; ++q)
        orc_ibf_bulk_count(words,
                           g,
                           values + offsets[q],
                           offsets[q + 1] - offsets[q],
                           width,
                           (char *)counts + (size_t)q * (size_t)g->bins * (size_t)width);
}

/* interleaved_bloom_filter.hpp:744-764: uint16 counts, then every bin with
 * count >= threshold in ascending order.  Returns the number of bins written. */
uint64_t orc_ibf_membership_for(uint64_t const * words,
                                orc_ibf const * g,
                                uint64_t const * values,
                                uint64_t n,
                                uint16_t threshold,
                                uint64_t * out)
{
    uint16_t * counts = (uint16_t *)malloc(sizeof(uint16_t) * (g->bins ? g->bins : 1));
    uint64_t found = 0;
    orc_ibf_bulk_count(words, g, values, n, 2, counts);
    for (uint64_t b = 0; b < g->bins; ++b)
        if (counts[b] >= threshold)
            out[found++] = b;
    free(counts);
    return found;
}

/* An HIBF as flat arrays: per-IBF geometry and word offset into one arena;
 * per-IBF technical-bin tables in CSR form (tb_offsets[i] .. tb_offsets[i+1]),
 * holding next_ibf_id and ibf_bin_to_user_bin_id
 * (hierarchical_interleaved_bloom_filter.hpp:209-238). */
typedef struct orc_hibf
{
    uint64_t n_ibfs;
    orc_ibf const * ibfs;
    uint64_t const * arena;
    uint64_t const * word_offsets;
    uint64_t const * tb_offsets;
    uint64_t const * next_ibf;
    uint64_t const * user_bin;
} orc_hibf;

typedef struct orc_out
{
    uint64_t * ids;
    uint64_t n;
    uint64_t cap;
} orc_out;

/* hierarchical_interleaved_bloom_filter.hpp:321-355 (membership_for_impl):
 * pre-order DFS, uint16 per-bin counts and a uint16 running sum.  Merged bin:
 * recurse iff sum >= t, then reset.  Regular bin at the end of its run (last
 * bin, or the next bin has another user-bin id): report iff sum >= t, then
 * reset.  Deleted bin: neither report nor reset. */
static void orc_hibf_membership_impl(orc_hibf const * hb,
                                     uint64_t const * values,
                                     uint64_t n,
                                     uint64_t ibf_idx,
                                     uint16_t threshold,
                                     orc_out * out)
{
    orc_ibf const * g = &hb->ibfs[ibf_idx];
    uint16_t * counts = (uint16_t *)malloc(sizeof(uint16_t) * (g->bins ? g->bins : 1));
    uint64_t const * ub = hb->user_bin + hb->tb_offsets[ibf_idx];
    uint64_t const * nx = hb->next_ibf + hb->tb_offsets[ibf_idx];
    uint16_t sum = 0;
    orc_ibf_bulk_count(hb->arena + hb->word_offsets[ibf_idx], g, values, n, 2, counts);
    for (uint64_t bin = 0; bin < g->bins; ++bin)
    {
        sum = (uint16_t)(sum + counts[bin]);
        uint64_t const id = ub[bin];
        if (id == ORC_BIN_DELETED)
            continue;
        if (id == ORC_BIN_MERGED)
        {
            if (sum >= threshold)
                orc_hibf_membership_impl(hb, values, n, nx[bin], threshold, out);
            sum = 0;
            continue;
        }
        if (bin + 1u == g->bins || id != ub[bin + 1u])
        {
            if (sum >= threshold)
            {
                if (out->n < out->cap)
                    out->ids[out->n] = id;
                ++out->n;
            }
            sum = 0;
        }
    }
    free(counts);
}

/* hierarchical_interleaved_bloom_filter.hpp:424-439.  Writes at most cap ids
 * in the reference's (unsorted, DFS) order and returns how many there are. */
uint64_t orc_hibf_membership_for(orc_hibf const * hb,
                                 uint64_t const * values,
                                 uint64_t n,
                                 uint16_t threshold,
                                 uint64_t * out,
                                 uint64_t cap)
{
    orc_out o = {out, 0, cap};
    orc_hibf_membership_impl(hb, values, n, 0, threshold, &o);
    return o.n;
}

/* Batched membership over CSR queries.  out_offsets[q] .. out_offsets[q+1]
 * index the ids of query q; out_cap bounds the total.  Returns the total
 * number of ids (which may exceed out_cap, in which case ids are truncated). */
uint64_t orc_hibf_membership_batch(orc_hibf const * hb,
                                   uint64_t const * values,
                                   uint64_t const * offsets,
                                   uint64_t nq,
                                   uint16_t const * thresholds,
                                   uint64_t * out_ids,
                                   uint64_t out_cap,
                                   uint64_t * out_offsets)
{
    uint64_t total = 0;
    out_offsets[0] = 0;
    for (uint64_t q = 0; q < nq; ++q)
    {
        uint64_t cap = total < out_cap ? out_cap - total : 0;
        uint64_t got = orc_hibf_membership_for(hb,
                                               values + offsets[q],
                                               offsets[q + 1] - offsets[q],
                                               thresholds[q],
                                               out_ids + (total < out_cap ? total : out_cap),
                                               cap);
        total += got;
        out_offsets[q + 1] = total;
    }
    return total;
}

/* hierarchical_interleaved_bloom_filter.hpp:462-496 (bulk_count_impl): the same
 * descent with value_t (width bytes, unsigned) counts and sum and a size_t
 * threshold; counts[ub] = sum iff sum >= threshold. */
static void orc_hibf_count_impl(orc_hibf const * hb,
                                uint64_t const * values,
                                uint64_t n,
                                uint64_t ibf_idx,
                                uint64_t threshold,
                                int width,
                                void * result)
{
    orc_ibf const * g = &hb->ibfs[ibf_idx];
    void * counts = malloc((size_t)width * (g->bins ? g->bins : 1));
    uint64_t const * ub = hb->user_bin + hb->tb_offsets[ibf_idx];
    uint64_t const * nx = hb->next_ibf + hb->tb_offsets[ibf_idx];
    uint64_t const mask = orc_width_mask(width);
    uint64_t sum = 0;
    orc_ibf_bulk_count(hb->arena + hb->word_offsets[ibf_idx], g, values, n, width, counts);
    for (uint64_t bin = 0; bin < g->bins; ++bin)
    {
        sum = (sum + orc_load(counts, bin, width)) & mask;
        uint64_t const id = ub[bin];
        if (id == ORC_BIN_DELETED)
            continue;
        if (id == ORC_BIN_MERGED)
        {
            if (sum >= threshold)
                orc_hibf_count_impl(hb, values, n, nx[bin], threshold, width, result);
            sum = 0;
            continue;
        }
        if (bin + 1u == g->bins || id != ub[bin + 1u])
        {
            if (sum >= threshold)
                orc_store(result, id, width, sum);
            sum = 0;
        }
    }
    free(counts);
}

/* hierarchical_interleaved_bloom_filter.hpp:549-566: zero n_user_bins counters,
 * then descend from IBF 0. */
void orc_hibf_bulk_count(orc_hibf const * hb,
                         uint64_t n_user_bins,
                         uint64_t const * values,
                         uint64_t n,
                         uint64_t threshold,
                         int width,
                         void * result)
{
    memset(result, 0, (size_t)n_user_bins * (size_t)width);
    orc_hibf_count_impl(hb, values, n, 0, threshold, width, result);
}

void orc_hibf_bulk_count_batch(orc_hibf const * hb,
                               uint64_t n_user_bins,
                               uint64_t const * values,
                               uint64_t const * offsets,
                               uint64_t nq,
                               uint64_t const * thresholds,
                               int width,
                               void * result,
                               int threads)
{
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads > 0 ? threads : 1)
#endif
    for (int64_t q = 0; q < (int64_t)nq; ++q)
        orc_hibf_bulk_count(hb,
                            n_user_bins,
                            values + offsets[q],
                            offsets[q + 1] - offsets[q],
                            thresholds[q],
                            width,
                            (char *)result + (size_t)q * (size_t)n_user_bins * (size_t)width);
}

/* Multi-threaded membership for the CPU baseline: per-thread agents over a
 * static partition of the queries; only the per-query hit counts are kept. */
void orc_hibf_membership_count_batch(orc_hibf const * hb,
                                     uint64_t const * values,
                                     uint64_t const * offsets,
                                     uint64_t nq,
                                     uint16_t const * thresholds,
                                     uint64_t * hits_per_query,
                                     int threads)
{
    (void)threads;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads > 0 ? threads : 1)
#endif
    {
        uint64_t scratch[1];
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int64_t q = 0; q < (int64_t)nq; ++q)
            hits_per_query[q] = orc_hibf_membership_for(hb,
                                                        values + offsets[q],
                                                        offsets[q + 1] - offsets[q],
                                                        thresholds[q],
                                                        scratch,
                                                        0);
    }
}

/* ------------------------------------------------------------------------
 * HyperLogLog sketches (src/sketch/hyperloglog.cpp), as the reference's build
 * uses them for user-bin cardinalities (compute_sketches.cpp:27-43,
 * estimate_kmer_counts.hpp).  Pinned by the reference's hyperloglog_test.cpp
 * expectations (the 8-value estimate 7.899522493 at 5 bits with the stated
 * index/rank of each value, the empty estimate 0, merge == add, the 1,500
 * mt19937_64 values within the error bound) and, for the 2^bits >= 128 branch
 * of alpha, by the mean relative error of many sketches against exact counts.
 * ------------------------------------------------------------------------ */

/* hyperloglog.cpp:48-54: 128-bit product with the golden-ratio constant, low
 * half xor high half. */
uint64_t orc_hll_wyhash(uint64_t value)
{
    unsigned __int128 const r = (unsigned __int128)value * 0x9E3779B97F4A7C15ULL;
    return (uint64_t)r ^ (uint64_t)(r >> 64);
}

/* hyperloglog.cpp:56-64: register index = top `bits` bits of the hash; rank =
 * leading zeros of the remaining bits (capped at 64 - bits) + 1; keep the max. */
void orc_hll_add(uint8_t * regs, uint32_t bits, uint64_t const * values, uint64_t n)
{
    uint64_t const rank_mask = (bits >= 64u) ? ~0ULL : ((1ULL << bits) - 1u);
    for (uint64_t i = 0; i < n; ++i)
    {
        uint64_t const h = orc_hll_wyhash(values[i]);
        uint64_t const idx = h >> (64u - bits);
        uint64_t const rest = (h << bits) | rank_mask;
        uint8_t const rank = (uint8_t)(__builtin_clzll(rest) + 1);
        if (rank > regs[idx])
            regs[idx] = rank;
    }
}

/* hyperloglog.cpp:136-149: register-wise max. */
void orc_hll_merge(uint8_t * regs, uint8_t const * other, uint32_t bits)
{
    uint64_t const size = 1ULL << bits;
    for (uint64_t i = 0; i < size; ++i)
        if (other[i] > regs[i])
            regs[i] = other[i];
}

/* hyperloglog.cpp:66-129.  The float sum runs in 8 lanes: for each 32-register
 * block and j in {0, 8, 16, 24}, lane L adds 2^-reg[j + 7 - L] (the argument
 * order of _mm256_set_ps puts register j + 7 in lane 0); the lanes are then
 * added in order 0..7.  alpha: 0.697 (32 registers), 0.709 (64), otherwise
 * 0.7213 / (1 + 1.079 / m) (hyperloglog.cpp:31-46).  Small-value correction:
 * m * ln(m / zeros) when the raw estimate is <= 2.5 m and a register is 0. */
double orc_hll_estimate(uint8_t const * regs, uint32_t bits)
{
    uint64_t const size = 1ULL << bits;
    double alpha;
    if (size == 32u)
        alpha = 0.697;
    else if (size == 64u)
        alpha = 0.709;
    else
        alpha = 0.7213 / (1.0 + 1.079 / (double)size);
    double const norm = alpha * (double)size * (double)size;
    float lane[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint64_t b = 0; b < size / 32u; ++b)
    {
        uint8_t const * const cur = regs + 32u * b;
        for (uint32_t j = 0; j <= 24u; j += 8u)
            for (uint32_t L = 0; L < 8u; ++L)
                lane[L] = lane[L] + 1.0f / (float)(1ULL << cur[j + 7u - L]);
    }
    float sum = 0.0f;
    for (uint32_t L = 0; L < 8u; ++L)
        sum = sum + lane[L];
    double estimate = norm / (double)sum;
    if (estimate <= 2.5 * (double)size)
    {
        uint32_t zeros = 0;
        for (uint64_t i = 0; i < size; ++i)
            zeros += regs[i] == 0u;
        if (zeros != 0u)
            estimate = (double)size * log((double)size / (double)zeros);
    }
    return estimate;
}

/* The layout DP's union estimates (src/sketch/toolbox.cpp:35-51,
 * precompute_union_estimates_for): for every column j < np and k < j, the
 * estimate of the union of the sketches at positions[k .. j], merged one by
 * one downward from j and truncated to uint64.  Column j's j values start at
 * out[j (j - 1) / 2].  `regs` holds the sketches back to back; `scratch` one
 * sketch. */
void orc_hll_union_estimates(uint8_t const * regs, uint32_t bits, uint64_t const * positions, uint64_t np,
                             uint8_t * scratch, uint64_t * out)
{
    uint64_t const size = 1ULL << bits;
    for (uint64_t j = 1; j < np; ++j)
    {
        memcpy(scratch, regs + positions[j] * size, size);
        for (uint64_t k = j; k-- > 0;)
        {
            orc_hll_merge(scratch, regs + positions[k] * size, bits);
            out[j * (j - 1u) / 2u + k] = (uint64_t)orc_hll_estimate(scratch, bits);
        }
    }
}

/* std::mt19937_64 (the C++ standard's parameters, [rand.predef]) seeded with
 * `seed`; the reference's HLL tests draw their inputs from it through
 * uniform_int_distribution<uint64_t> over the full range, which returns the
 * engine's outputs unchanged. */
void orc_mt19937_64(uint64_t seed, uint64_t * out, uint64_t n)
{
    enum { NN = 312, MM = 156 };
    uint64_t mt[NN];
    uint64_t const matrix_a = 0xB5026F5AA96619E9ULL;
    uint64_t const upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
    mt[0] = seed;
    for (int i = 1; i < NN; ++i)
        mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
    int idx = NN;
    for (uint64_t k = 0; k < n; ++k)
    {
        if (idx >= NN)
        {
            for (int i = 0; i < NN; ++i)
            {
                uint64_t const x = (mt[i] & upper) | (mt[(i + 1) % NN] & lower);
                mt[i] = mt[(i + MM) % NN] ^ (x >> 1) ^ ((x & 1u) ? matrix_a : 0u);
            }
            idx = 0;
        }
        uint64_t y = mt[idx++];
        y ^= (y >> 29) & 0x5555555555555555ULL;
        y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
        y ^= (y << 37) & 0xFFF7EEE000000000ULL;
        y ^= y >> 43;
        out[k] = y;
    }
}

/* ------------------------------------------------------------------------
 * Split-bin chunk order.  The reference chunks a split user bin in the
 * iteration order of a robin_hood::unordered_flat_set<uint64_t>
 * (src/build/insert_into_ibf.cpp:28-47; robin_hood 3.11.5 vendored as
 * include/hibf/contrib/robin_hood.hpp, Table<true, 80, uint64_t, void, ...>).
 * orc_fs restates that table's slot order: keyToIdx (hash_int, then
 * h *= multiplier, h ^= h >> 33; info = inc + ((h & 31) >> shift), home =
 * (h >> 5) & mask), insertKeyPrepareEmptySpot, shiftUp, insert_move,
 * rehashPowerOfTwo, initData, try_increase_info, increase_size and clear.
 * Iteration is slot order.  Pinned by the golden orders the reference's own
 * set writes (oracle/ref_robin_hood.cpp -> tests/golden/robin_hood_orders.json)
 * and, where the reference tree is present, by running that program.
 * ------------------------------------------------------------------------ */
typedef struct orc_fs
{
    uint64_t * keys;
    uint8_t * info; /* nb slots, sentinel, 7 padding bytes */
    uint64_t mask, n, max_allowed, multiplier;
    uint32_t inc, shift;
} orc_fs;

static uint64_t orc_fs_max_allowed(uint64_t slots)
{
    return slots <= UINT64_MAX / 100u ? slots * 80u / 100u : (slots / 100u) * 80u;
}

static uint64_t orc_fs_with_buffer(uint64_t slots)
{
    uint64_t const m = orc_fs_max_allowed(slots);
    return slots + (m < 255u ? m : 255u);
}

void orc_fs_init(orc_fs * s)
{
    memset(s, 0, sizeof(*s));
    s->multiplier = 0xC4CEB9FE1A85EC53ULL;
    s->inc = 32u;
}

void orc_fs_free(orc_fs * s)
{
    free(s->keys);
    free(s->info);
    orc_fs_init(s);
}

static void orc_fs_key_to_idx(orc_fs const * s, uint64_t key, uint64_t * idx, uint32_t * info)
{
    uint64_t h = key;
    h ^= h >> 33;
    h *= 0xFF51AFD7ED558CCDULL; /* robin_hood::hash_int */
    h ^= h >> 33;
    h *= s->multiplier;
    h ^= h >> 33;
    *info = s->inc + (uint32_t)((h & 31u) >> s->shift);
    *idx = (h >> 5) & s->mask;
}

static void orc_fs_init_data(orc_fs * s, uint64_t slots)
{
    uint64_t const nb = orc_fs_with_buffer(slots);
    s->n = 0;
    s->mask = slots - 1u;
    s->max_allowed = orc_fs_max_allowed(slots);
    s->keys = (uint64_t *)malloc(nb * sizeof(uint64_t));
    s->info = (uint8_t *)calloc(nb + 8u, 1);
    s->info[nb] = 1;
    s->inc = 32u;
    s->shift = 0;
}

static int orc_fs_try_increase_info(orc_fs * s)
{
    uint64_t nb;
    if (s->inc <= 2u)
        return 0;
    s->inc >>= 1;
    ++s->shift;
    nb = orc_fs_with_buffer(s->mask + 1u);
    for (uint64_t i = 0; i < nb; i += 8)
    {
        uint64_t v;
        memcpy(&v, s->info + i, 8);
        v = (v >> 1) & 0x7F7F7F7F7F7F7F7FULL;
        memcpy(s->info + i, &v, 8);
    }
    s->info[nb] = 1;
    s->max_allowed = orc_fs_max_allowed(s->mask + 1u);
    return 1;
}

static void orc_fs_shift_up(orc_fs * s, uint64_t start, uint64_t insertion_idx)
{
    uint64_t idx = start;
    s->keys[idx] = s->keys[idx - 1];
    while (--idx != insertion_idx)
        s->keys[idx] = s->keys[idx - 1];
    idx = start;
    while (idx != insertion_idx)
    {
        s->info[idx] = (uint8_t)(s->info[idx - 1] + s->inc);
        if (s->info[idx] + s->inc > 0xFFu)
            s->max_allowed = 0;
        --idx;
    }
}

static void orc_fs_insert_move(orc_fs * s, uint64_t key)
{
    uint64_t idx, insertion_idx;
    uint32_t info;
    uint8_t insertion_info;
    if (s->max_allowed == 0 && !orc_fs_try_increase_info(s))
        abort(); /* the set throws an overflow error here */
    orc_fs_key_to_idx(s, key, &idx, &info);
    while (info <= s->info[idx])
    {
        ++idx;
        info += s->inc;
    }
    insertion_idx = idx;
    insertion_info = (uint8_t)info;
    if (insertion_info + s->inc > 0xFFu)
        s->max_allowed = 0;
    while (s->info[idx] != 0)
    {
        ++idx;
        info += s->inc;
    }
    if (idx != insertion_idx)
        orc_fs_shift_up(s, idx, insertion_idx);
    s->keys[insertion_idx] = key;
    s->info[insertion_idx] = insertion_info;
    ++s->n;
}

static void orc_fs_rehash(orc_fs * s, uint64_t slots)
{
    uint64_t * const old_keys = s->keys;
    uint8_t * const old_info = s->info;
    uint64_t const old_nb = orc_fs_with_buffer(s->mask + 1u);
    orc_fs_init_data(s, slots);
    if (old_info && old_nb > 1u)
        for (uint64_t i = 0; i < old_nb; ++i)
            if (old_info[i] != 0)
                orc_fs_insert_move(s, old_keys[i]);
    free(old_keys);
    free(old_info);
}

static void orc_fs_increase_size(orc_fs * s)
{
    if (s->mask == 0)
    {
        orc_fs_init_data(s, 8u);
        return;
    }
    if (s->n < orc_fs_max_allowed(s->mask + 1u) && orc_fs_try_increase_info(s))
        return;
    if (s->n * 2u < orc_fs_max_allowed(s->mask + 1u))
    {
        s->multiplier += 0xC4CEB9FE1A85EC54ULL;
        orc_fs_rehash(s, s->mask + 1u);
    }
    else
    {
        orc_fs_rehash(s, (s->mask + 1u) * 2u);
    }
}

/* emplace(key): 1 when inserted, 0 when present. */
int orc_fs_emplace(orc_fs * s, uint64_t key)
{
    for (int attempt = 0; attempt < 256; ++attempt)
    {
        uint64_t idx, insertion_idx;
        uint32_t info, insertion_info;
        orc_fs_key_to_idx(s, key, &idx, &info);
        if (s->mask != 0)
        {
            while (info < s->info[idx])
            {
                ++idx;
                info += s->inc;
            }
            while (info == s->info[idx])
            {
                if (s->keys[idx] == key)
                    return 0;
                ++idx;
                info += s->inc;
            }
        }
        if (s->n >= s->max_allowed)
        {
            orc_fs_increase_size(s);
            continue;
        }
        insertion_idx = idx;
        insertion_info = info;
        if (insertion_info + s->inc > 0xFFu)
            s->max_allowed = 0;
        while (s->info[idx] != 0)
        {
            ++idx;
            info += s->inc;
        }
        if (idx != insertion_idx)
            orc_fs_shift_up(s, idx, insertion_idx);
        s->keys[insertion_idx] = key;
        s->info[insertion_idx] = (uint8_t)insertion_info;
        ++s->n;
        return 1;
    }
    abort(); /* the set throws an overflow error here */
}

void orc_fs_clear(orc_fs * s)
{
    uint64_t nb;
    if (s->n == 0)
        return;
    s->n = 0;
    nb = orc_fs_with_buffer(s->mask + 1u);
    memset(s->info, 0, nb + 8u);
    s->info[nb] = 1;
    s->inc = 32u;
    s->shift = 0;
}

/* Writes the keys in iteration order to out (s->n entries). */
void orc_fs_order(orc_fs const * s, uint64_t * out)
{
    uint64_t k = 0;
    if (s->mask == 0)
        return;
    for (uint64_t i = 0, nb = orc_fs_with_buffer(s->mask + 1u); i < nb; ++i)
        if (s->info[i] != 0)
            out[k++] = s->keys[i];
}

/* The golden-order script protocol of oracle/ref_robin_hood.cpp, run on
 * orc_fs: sets 0..63.  Appends each dump line to out (NUL-terminated);
 * returns the bytes needed (more than cap: out is cut). */
uint64_t orc_fs_script(char const * script, char * out, uint64_t cap)
{
    orc_fs sets[64];
    uint64_t used = 0;
    char const * p = script;
    for (int i = 0; i < 64; ++i)
        orc_fs_init(&sets[i]);
    if (cap)
        out[0] = 0;
    while (*p)
    {
        char op[16] = {0};
        int a = 0, consumed = 0;
        char const * eol = strchr(p, '\n');
        uint64_t const len = eol ? (uint64_t)(eol - p) : strlen(p);
        char * const line = (char *)malloc(len + 1u);
        memcpy(line, p, len);
        line[len] = 0;
        p += len + (eol ? 1u : 0u);
        if (sscanf(line, "%15s %d%n", op, &a, &consumed) < 2 || a < 0 || a >= 64)
        {
            free(line);
            continue;
        }
        {
            orc_fs * const s = &sets[a];
            char const * rest = line + consumed;
            if (!strcmp(op, "gen"))
            {
                unsigned long long seed, set, begin, count;
                if (sscanf(rest, "%llu %llu %llu %llu", &seed, &set, &begin, &count) == 4)
                    for (uint64_t j = 0; j < count; ++j)
                        orc_fs_emplace(s, orc_synth_value(seed, set, begin + j));
            }
            else if (!strcmp(op, "genmod"))
            {
                unsigned long long seed, set, count, mod;
                if (sscanf(rest, "%llu %llu %llu %llu", &seed, &set, &count, &mod) == 4 && mod)
                    for (uint64_t j = 0; j < count; ++j)
                        orc_fs_emplace(s, orc_synth_value(seed, set, j % mod));
            }
            else if (!strcmp(op, "vals"))
            {
                unsigned long long v;
                int used_chars = 0;
                while (sscanf(rest, "%llu%n", &v, &used_chars) == 1)
                {
                    orc_fs_emplace(s, v);
                    rest += used_chars;
                }
            }
            else if (!strcmp(op, "merge"))
            {
                int b = 0;
                if (sscanf(rest, "%d", &b) == 1 && b >= 0 && b < 64 && b != a)
                {
                    uint64_t * const order = (uint64_t *)malloc((sets[b].n + 1u) * sizeof(uint64_t));
                    uint64_t const n = sets[b].n;
                    orc_fs_order(&sets[b], order);
                    for (uint64_t j = 0; j < n; ++j)
                        orc_fs_emplace(s, order[j]);
                    free(order);
                }
            }
            else if (!strcmp(op, "clear"))
            {
                orc_fs_clear(s);
            }
            else if (!strcmp(op, "dump"))
            {
                int full = 0;
                uint64_t * const order = (uint64_t *)malloc((s->n + 1u) * sizeof(uint64_t));
                uint64_t h = 1469598103934665603ULL;
                char buf[64];
                int w;
                sscanf(rest, "%d", &full);
                orc_fs_order(s, order);
                for (uint64_t j = 0; j < s->n; ++j)
                {
                    h ^= order[j];
                    h *= 1099511628211ULL;
                }
                w = snprintf(buf, sizeof buf, "dump %d %llu %llu %llu", a, (unsigned long long)s->n,
                             (unsigned long long)s->mask, (unsigned long long)h);
                if (used + (uint64_t)w < cap)
                    memcpy(out + used, buf, (size_t)w + 1u);
                used += (uint64_t)w;
                if (full)
                    for (uint64_t j = 0; j < s->n; ++j)
                    {
                        w = snprintf(buf, sizeof buf, " %llu", (unsigned long long)order[j]);
                        if (used + (uint64_t)w < cap)
                            memcpy(out + used, buf, (size_t)w + 1u);
                        used += (uint64_t)w;
                    }
                if (used + 1u < cap)
                    memcpy(out + used, "\n", 2);
                used += 1u;
                free(order);
            }
        }
        free(line);
    }
    for (int i = 0; i < 64; ++i)
        orc_fs_free(&sets[i]);
    return used + 1u;
}

/* Heap-held sets for callers that keep several (the tests' restatement of the
 * reference's build replays one per IBF). */
orc_fs * orc_fs_new(void)
{
    orc_fs * s = (orc_fs *)malloc(sizeof(orc_fs));
    orc_fs_init(s);
    return s;
}

void orc_fs_delete(orc_fs * s)
{
    if (!s)
        return;
    orc_fs_free(s);
    free(s);
}

uint64_t orc_fs_size(orc_fs const * s) { return s->n; }

/* compute_kmers over generated values: synth(seed, set, begin + j), j < count. */
void orc_fs_emplace_synth(orc_fs * s, uint64_t seed, uint64_t set, uint64_t begin, uint64_t count)
{
    for (uint64_t j = 0; j < count; ++j)
        orc_fs_emplace(s, orc_synth_value(seed, set, begin + j));
}

void orc_fs_emplace_values(orc_fs * s, uint64_t const * values, uint64_t n)
{
    for (uint64_t j = 0; j < n; ++j)
        orc_fs_emplace(s, values[j]);
}

/* update_parent_kmers: dst.insert(src.begin(), src.end()). */
void orc_fs_insert_set(orc_fs * dst, orc_fs const * src)
{
    uint64_t * const order = (uint64_t *)malloc((src->n + 1u) * sizeof(uint64_t));
    uint64_t const n = src->n;
    orc_fs_order(src, order);
    for (uint64_t j = 0; j < n; ++j)
        orc_fs_emplace(dst, order[j]);
    free(order);
}

/* ------------------------------------------------------------------------
 * Minimiser hashes of a DNA read: seqan3::views::minimiser_hash(shape,
 * window_size, seed) over dna4, seqan3 3.4.2 (the version the reference pins
 * for its tutorial, cmake/package-lock.cmake:50-56; seqan3 is not part of the
 * reference tree, and the reference's own tutorial hashes with kmer_hash,
 * test/tutorial/search.cpp:40).  Restated from seqan3's published definition
 * (search/views/minimiser_hash.hpp, search/views/minimiser.hpp):
 *   - shape: bits as written in seqan3's binary literal, the most significant
 *     set bit is the window's first position (0b1101: use, use, skip, use);
 *     size = bit length, the k-mer hash is the base-4 number of the dna4 ranks
 *     at the used positions, first position most significant;
 *   - forward[i]  = hash of the window at i, XOR seed;
 *   - reverse[i]  = the hash at i of complement(text) reversed, taken from the
 *     far end, XOR seed (the shape is applied to the reverse complement
 *     window, not mirrored);
 *   - value[i] = min(forward[i], reverse[i]); n = len - size + 1 values;
 *   - a window of W = window_size - size + 1 values (W = n when n < W); the
 *     first window emits its minimum, the rightmost one on ties; after that,
 *     when a value enters, the tracked minimum is recomputed (rightmost on
 *     ties) and emitted if its position just left the window, else the
 *     entering value is emitted and tracked if strictly smaller, else nothing.
 * Pinned by: the chain form of the same rule computed independently in
 * tests/test_oracle.py (next position = the first strictly smaller value
 * within W, else the rightmost minimum of the following W values), W = 1
 * (every canonical value is emitted), the forward hashes against the k-mer
 * oracle, and window-minimum invariants.  Parity with seqan3's own output is
 * unpinned (no seqan3 output is held anywhere the tests can read). */
static inline uint64_t orc_dna4_rank(uint8_t c)
{
    switch (c | 0x20u)
    {
    case 'c': return 1u;
    case 'g': return 2u;
    case 't':
    case 'u': return 3u;
    default: return 0u;
    }
}

static uint64_t orc_shape_hash(uint8_t const * seq, uint64_t i, uint64_t shape, uint32_t size, int reverse)
{
    uint64_t h = 0;
    for (uint32_t t = 0; t < size; ++t)
    {
        if (!((shape >> (size - 1u - t)) & 1u))
            continue;
        uint64_t const r = reverse ? 3u - orc_dna4_rank(seq[i + size - 1u - t]) : orc_dna4_rank(seq[i + t]);
        h = (h << 2) | r;
    }
    return h;
}

/* Writes up to cap hashes to out; returns the number of minimisers (also when
 * it exceeds cap).  shape != 0 with bit length <= 32 and window_size >= that
 * length are the caller's to check. */
uint64_t orc_minimiser_hash(uint8_t const * seq, uint64_t len, uint64_t shape, uint32_t window_size, uint64_t seed,
                            uint64_t * out, uint64_t cap)
{
    uint32_t size = 0;
    while (size < 64u && (shape >> size))
        ++size;
    if (len < size)
        return 0;
    uint64_t const n = len - size + 1u;
    uint64_t w = (uint64_t)window_size - size + 1u;
    if (w > n)
        w = n;
    uint64_t * const ring = (uint64_t *)malloc(w * sizeof(uint64_t)); /* window values, oldest at head */
    uint64_t head = 0, count = 0;
    uint64_t min_value = 0, min_offset = 0; /* offset from the window's oldest value */
#define ORC_VALUE(i) ({ uint64_t f_ = orc_shape_hash(seq, (i), shape, size, 0) ^ seed; \
                        uint64_t r_ = orc_shape_hash(seq, (i), shape, size, 1) ^ seed; f_ < r_ ? f_ : r_; })
#define ORC_EMIT(v) do { if (count < cap) out[count] = (v); ++count; } while (0)
    for (uint64_t i = 0; i < w; ++i)
        ring[i] = ORC_VALUE(i);
    for (uint64_t i = 0; i < w; ++i)
        if (i == 0 || ring[i] <= min_value)
        {
            min_value = ring[i];
            min_offset = i;
        }
    ORC_EMIT(min_value);
    for (uint64_t i = w; i < n; ++i)
    {
        uint64_t const v = ORC_VALUE(i);
        ring[head] = v; /* the oldest value leaves, v enters at the back */
        head = (head + 1u) % w;
        if (min_offset == 0)
        {
            for (uint64_t j = 0; j < w; ++j)
            {
                uint64_t const x = ring[(head + j) % w];
                if (j == 0 || x <= min_value)
                {
                    min_value = x;
                    min_offset = j;
                }
            }
            ORC_EMIT(min_value);
        }
        else if (v < min_value)
        {
            min_value = v;
            min_offset = w - 1u;
            ORC_EMIT(min_value);
        }
        else
            --min_offset;
    }
#undef ORC_VALUE
#undef ORC_EMIT
    free(ring);
    return count;
}
