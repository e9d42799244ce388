// Host-side layout computation: a restatement of the reference's
// hierarchical_binning / simple_binning DP (src/layout/hierarchical_binning.cpp,
// src/layout/simple_binning.cpp), with union estimates and similarity
// rearrangement when sketches are given, plus the layout text format
// (src/layout/layout.cpp).
#include "layout.hpp"

#include <algorithm>
#include <barrier>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <thread>
#include <limits>
#include <memory>
#include <sstream>
#include <utility>

namespace hibf_amd::host
{

std::string validate_and_set_defaults(Config & c)
{
    if (c.number_of_user_bins == 0u)
        return "[HIBF CONFIG ERROR] You did not set the required config::number_of_user_bins.";
    if (c.number_of_user_bins >= 18'446'744'073'709'551'614ULL)
        return "[HIBF CONFIG ERROR] The maximum possible config::number_of_user_bins is 18446744073709551613.";
    if (c.number_of_hash_functions == 0u || c.number_of_hash_functions > 5u)
        return "[HIBF CONFIG ERROR] config::number_of_hash_functions must be in [1,5].";
    if (c.maximum_fpr <= 0.0 || c.maximum_fpr >= 1.0)
        return "[HIBF CONFIG ERROR] config::maximum_fpr must be in (0.0,1.0).";
    if (c.relaxed_fpr <= 0.0 || c.relaxed_fpr >= 1.0)
        return "[HIBF CONFIG ERROR] config::relaxed_fpr must be in (0.0,1.0).";
    if (c.relaxed_fpr < c.maximum_fpr)
        return "[HIBF CONFIG ERROR] config::relaxed_fpr must be greater than or equal to config::maximum_fpr.";
    if (c.threads == 0u)
        return "[HIBF CONFIG ERROR] config::threads must be greater than 0.";
    if (c.sketch_bits < 5u || c.sketch_bits > 32u)
        return "[HIBF CONFIG ERROR] config::sketch_bits must be in [5,32].";
    if (c.tmax == 0)
        c.tmax = next_multiple_of_64(static_cast<uint64_t>(std::ceil(std::sqrt(c.number_of_user_bins))));
    else if (c.tmax > 18'446'744'073'709'551'552ULL)
        return "[HIBF CONFIG ERROR] The maximum possible config::tmax is 18446744073709551552.";
    else if (c.tmax % 64 != 0)
        c.tmax = next_multiple_of_64(c.tmax);
    if (c.empty_bin_fraction < 0.0 || c.empty_bin_fraction >= 1.0)
        return "[HIBF CONFIG ERROR] config::empty_bin_fraction must be in [0.0,1.0).";
    if (c.empty_bin_fraction != 0.0)
        c.track_occupancy = true;
    if (c.alpha < 0.0)
        return "[HIBF CONFIG ERROR] config::alpha must be positive.";
    if (c.max_rearrangement_ratio < 0.0 || c.max_rearrangement_ratio > 1.0)
        return "[HIBF CONFIG ERROR] config::max_rearrangement_ratio must be in [0.0,1.0].";
    if (c.disable_estimate_union || c.max_rearrangement_ratio == 0.0)
        c.disable_rearrangement = true;
    return {};
}

uint64_t bin_size_in_bits(double fpr, uint64_t hash_count, uint64_t elements)
{
    double const numerator{-static_cast<double>(elements * hash_count)};
    double const denominator{std::log(1 - std::exp(std::log(fpr) / hash_count))};
    double const result{std::ceil(numerator / denominator)};
    return static_cast<uint64_t>(result);
}

std::vector<double> fpr_correction(double fpr, uint64_t hash_count, uint64_t t_max)
{
    uint64_t const max_tb = next_multiple_of_64(t_max);
    std::vector<double> corr(max_tb + 1u, 0.0);
    corr[1] = 1.0;
    double const numerator = std::log1p(-std::exp(std::log(fpr) / hash_count));
    for (uint64_t split = 2u; split <= max_tb; ++split)
    {
        double const log_target_fpr = std::log1p(-std::exp(std::log1p(-fpr) / split));
        corr[split] = numerator / std::log1p(-std::exp(log_target_fpr / hash_count));
    }
    return corr;
}

double relaxed_fpr_correction(double fpr, double relaxed_fpr, uint64_t hash_count)
{
    double const numerator = std::log1p(-std::exp(std::log(fpr) / hash_count));
    double const denominator = std::log1p(-std::exp(std::log(relaxed_fpr) / hash_count));
    return numerator / denominator;
}

uint64_t subtract_empty_bins(uint64_t tmax, double fraction)
{
    if (fraction == 0.0 || tmax <= 2u)
        return tmax;
    uint64_t const n = std::clamp<uint64_t>(static_cast<uint64_t>(tmax * fraction), 1, tmax - 2);
    return tmax - n;
}

namespace
{

uint64_t divide_and_ceil(uint64_t a, uint64_t b) { return (a + (b - 1u)) / b; }

struct Store
{
    Layout * layout;
    std::vector<uint64_t> const * kmer_counts;
    std::vector<uint64_t> positions;
    std::vector<double> const * fpr_correction;
    double relaxed_fpr_correction;
    std::vector<uint64_t> previous; // bin indices of the merged bins above
    Sketching * sketching;          // null: sums instead of union estimates
};

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct MaxTracker
{
    uint64_t max_id{};
    uint64_t max_size{};
    void update(uint64_t id, uint64_t size)
    {
        if (size > max_size)
        {
            max_id = id;
            max_size = size;
        }
    }
};

uint64_t simple_binning(Store & d, uint64_t num_technical_bins);
uint64_t hierarchical_binning(Store & d, Config const & c);

// hierarchical_binning.cpp:394-427 (add_lower_level)
uint64_t add_lower_level(Store & lower, Config const & c)
{
    uint64_t const n = lower.positions.size();
    if (n > c.tmax)
        return hierarchical_binning(lower, c);
    uint64_t const tbs = subtract_empty_bins(next_multiple_of_64(n), c.empty_bin_fraction);
    return simple_binning(lower, tbs);
}

// simple_binning.cpp:19-116
uint64_t simple_binning(Store & d, uint64_t num_technical_bins)
{
    auto const & counts = *d.kmer_counts;
    auto const & corr = *d.fpr_correction;
    uint64_t const num_user_bins = d.positions.size();
    uint64_t const maxv = std::numeric_limits<uint64_t>::max();
    std::vector<std::vector<uint64_t>> matrix(num_technical_bins, std::vector<uint64_t>(num_user_bins, maxv));
    std::vector<std::vector<uint64_t>> trace(num_technical_bins, std::vector<uint64_t>(num_user_bins, maxv));
    uint64_t const extra_bins = num_technical_bins - num_user_bins + 1;
    double const init_card = static_cast<double>(counts[d.positions[0]]);
    for (uint64_t i = 0; i < extra_bins; ++i)
    {
        uint64_t const corrected = static_cast<uint64_t>(init_card * corr[i + 1]);
        matrix[i][0] = divide_and_ceil(corrected, i + 1u);
    }
    // The size of a split into s bins depends on the column only: split[s]
    // (the reference recomputes it for every cell).
    std::vector<uint64_t> split(extra_bins + 1u, 0u);
    for (uint64_t j = 1; j < num_user_bins; ++j)
    {
        double const card = static_cast<double>(counts[d.positions[j]]);
        for (uint64_t sb = 1; sb <= extra_bins; ++sb)
            split[sb] = divide_and_ceil(static_cast<uint64_t>(card * corr[sb]), sb);
        for (uint64_t i = j; i < j + extra_bins; ++i)
        {
            uint64_t minimum = maxv;
            for (uint64_t ip = j - 1; ip < i; ++ip)
            {
                uint64_t const score = std::max<uint64_t>(split[i - ip], matrix[ip][j - 1]);
                if (score < minimum)
                {
                    trace[i][j] = ip;
                    minimum = score;
                }
            }
            matrix[i][j] = minimum;
        }
    }
    uint64_t trace_i = num_technical_bins - 1;
    uint64_t trace_j = num_user_bins - 1;
    uint64_t max_id{}, max_size{}, bin_id{};
    while (trace_j > 0)
    {
        uint64_t const next_i = trace[trace_i][trace_j];
        uint64_t const nb = trace_i - next_i;
        uint64_t const card = counts[d.positions[trace_j]];
        uint64_t const corrected = static_cast<uint64_t>(card * corr[nb]);
        uint64_t const per_bin = divide_and_ceil(corrected, nb);
        d.layout->user_bins.push_back({d.previous, bin_id, nb, d.positions[trace_j]});
        if (per_bin > max_size)
        {
            max_id = bin_id;
            max_size = per_bin;
        }
        bin_id += nb;
        trace_i = next_i;
        --trace_j;
    }
    ++trace_i;
    uint64_t const card = counts[d.positions[0]];
    uint64_t const corrected = static_cast<uint64_t>(card * corr[trace_i]);
    uint64_t const per_bin = divide_and_ceil(corrected, trace_i);
    d.layout->user_bins.push_back({d.previous, bin_id, trace_i, d.positions[0]});
    if (per_bin > max_size)
        max_id = bin_id;
    return max_id;
}

uint64_t needed_technical_bins(Config const & c, uint64_t n)
{
    return subtract_empty_bins(std::min<uint64_t>(next_multiple_of_64(n), c.tmax), c.empty_bin_fraction);
}

uint64_t max_merge_levels(Config const & c, uint64_t n)
{
    uint64_t const lower = needed_technical_bins(c, n);
    double const levels = std::log(n) / std::log(lower);
    return static_cast<uint64_t>(std::ceil(levels));
}

// hierarchical_binning.cpp:31-79 and :102-393 with union estimates = sums.
uint64_t hierarchical_binning(Store & d, Config const & c)
{
    auto const & counts = *d.kmer_counts;
    auto const & corr = *d.fpr_correction;
    double const rcorr = d.relaxed_fpr_correction;
    // hierarchical_binning.cpp:35-56: sort by count, then rearrange by similarity.
    seqan::hibf::sketch::toolbox::sort_by_cardinalities(counts, d.positions);
    Sketching * const sk = d.sketching;
    if (sk && !c.disable_rearrangement)
    {
        auto const t0 = std::chrono::steady_clock::now();
        if (sk->pair_unions)
            seqan::hibf::sketch::toolbox::rearrange_bins(*sk->sketches, counts, d.positions, c.max_rearrangement_ratio,
                                                         c.threads, *sk->pair_unions);
        else
            seqan::hibf::sketch::toolbox::rearrange_bins(*sk->sketches, counts, d.positions, c.max_rearrangement_ratio,
                                                         c.threads);
        sk->rearrangement_seconds += seconds_since(t0);
    }
    std::unique_ptr<HostUnionEstimates> host_unions;
    UnionEstimates * unions = nullptr;
    if (sk)
    {
        if (!sk->unions)
            host_unions = std::make_unique<HostUnionEstimates>(*sk->sketches);
        unions = sk->unions ? sk->unions : host_unions.get();
        auto const t0 = std::chrono::steady_clock::now();
        if (!unions->start(d.positions))
            throw std::runtime_error("union estimates failed");
        sk->union_seconds += seconds_since(t0);
    }

    uint64_t const num_user_bins = d.positions.size();
    uint64_t const num_technical_bins =
        d.previous.empty() ? subtract_empty_bins(c.tmax, c.empty_bin_fraction) : needed_technical_bins(c, num_user_bins);
    uint64_t const maxv = std::numeric_limits<uint64_t>::max();
    std::vector<std::vector<uint64_t>> matrix(num_technical_bins, std::vector<uint64_t>(num_user_bins, maxv));
    std::vector<std::vector<uint64_t>> ll(num_technical_bins, std::vector<uint64_t>(num_user_bins, 0u));
    std::vector<std::vector<std::pair<uint64_t, uint64_t>>> trace(
        num_technical_bins, std::vector<std::pair<uint64_t, uint64_t>>(num_user_bins, {maxv, maxv}));

    // initialization (:81-134)
    double const card0 = static_cast<double>(counts[d.positions[0]]);
    for (uint64_t i = 0; i < num_technical_bins; ++i)
    {
        uint64_t const corrected = static_cast<uint64_t>(card0 * corr[i + 1]);
        matrix[i][0] = divide_and_ceil(corrected, i + 1u);
        trace[i][0] = {0u, 0u};
    }
    uint64_t sum = counts[d.positions[0]];
    std::vector<uint64_t> union_est; // prefix unions (row 0 of the DP)
    uint64_t const * union_col = nullptr;
    if (unions)
    {
        union_est.assign(num_user_bins, 0u);
        auto const t0 = std::chrono::steady_clock::now();
        if (!unions->prefix(union_est.data()))
            throw std::runtime_error("union estimates failed");
        sk->union_seconds += seconds_since(t0);
    }
    for (uint64_t j = 1; j < num_user_bins; ++j)
    {
        sum += counts[d.positions[j]];
        matrix[0][j] = static_cast<uint64_t>((unions ? union_est[j] : sum) * rcorr);
        ll[0][j] = max_merge_levels(c, j + 1) * sum;
        trace[0][j] = {0u, j - 1};
    }

    // recursion (:136-232).  Cells of one column only read earlier columns, so
    // the rows of a column are computed in parallel (interleaved over worker
    // threads, a barrier per column); column j-1 is mirrored into contiguous
    // arrays for the i' scan.  The arithmetic per cell is the reference's.
    std::vector<uint64_t> mml(num_user_bins + 1, 0);
    for (uint64_t n = 1; n <= num_user_bins; ++n)
        mml[n] = max_merge_levels(c, n);
    std::vector<uint64_t> sorted_counts(num_user_bins);
    for (uint64_t j = 0; j < num_user_bins; ++j)
        sorted_counts[j] = counts[d.positions[j]];
    std::vector<uint64_t> colM(num_technical_bins), colL(num_technical_bins);
    for (uint64_t i = 0; i < num_technical_bins; ++i)
    {
        colM[i] = matrix[i][0];
        colL[i] = ll[i][0];
    }
    // split_size[s]: the size of column j's user bin split into s technical bins,
    // which the vertical scan of every cell of the column needs (computed once
    // per column here; the reference recomputes it per cell).
    std::vector<uint64_t> split_size(num_technical_bins, 0u);
    auto column_split = [&](uint64_t j)
    {
        double const card = static_cast<double>(sorted_counts[j]);
        for (uint64_t sb = 1; sb < num_technical_bins; ++sb)
            split_size[sb] = divide_and_ceil(static_cast<uint64_t>(card * corr[sb]), sb);
    };
    auto cell = [&](uint64_t i, uint64_t j)
    {
        uint64_t const current_weight = sorted_counts[j];
        uint64_t minimum = maxv;
        uint64_t full_minimum = maxv;
        std::pair<uint64_t, uint64_t> best{maxv, maxv};
        uint64_t best_ll = 0;
        for (uint64_t ip = 0; ip < i; ++ip)
        {
            uint64_t const score = std::max<uint64_t>(split_size[i - ip], colM[ip]);
            uint64_t const full_score =
                static_cast<uint64_t>(static_cast<double>(score * (i + 1)) + c.alpha * static_cast<double>(colL[ip]));
            if (full_score < full_minimum)
            {
                minimum = score;
                full_minimum = full_score;
                best = {ip, j - 1};
                best_ll = colL[ip];
            }
        }
        auto const & trow = trace[i];
        auto const & mprev = matrix[i - 1];
        auto const & lprev = ll[i - 1];
        uint64_t jp = j - 1;
        uint64_t weight = current_weight;
        // :191-198: the merged bin's size is the union estimate of {jp + 1 .. j}
        // (or the sum of the counts), relaxed-FPR corrected and truncated.
        auto get_weight = [&]() -> uint64_t
        {
            uint64_t const uncorrected = !unions ? weight : jp + 1 == j ? current_weight : union_col[jp + 1];
            return static_cast<uint64_t>(rcorr * static_cast<double>(uncorrected));
        };
        while (jp != 0 && ((i - trow[jp].first) < 2) && get_weight() < minimum)
        {
            weight += sorted_counts[jp];
            --jp;
            uint64_t const score = std::max<uint64_t>(mprev[jp], get_weight());
            uint64_t const ll_kmers = lprev[jp] + mml[j - jp] * weight;
            uint64_t const full_score =
                static_cast<uint64_t>(static_cast<double>(score * (i + 1)) + c.alpha * static_cast<double>(ll_kmers));
            if (full_score < full_minimum)
            {
                minimum = score;
                full_minimum = full_score;
                best = {i - 1, jp};
                best_ll = ll_kmers;
            }
        }
        trace[i][j] = best;
        ll[i][j] = best_ll;
        matrix[i][j] = minimum;
    };
    uint64_t const work = num_technical_bins * num_user_bins;
    unsigned const hw = std::max(1u, std::thread::hardware_concurrency());
    unsigned const nthreads = work < (1u << 16) ? 1u : std::min<unsigned>(16u, hw);
    auto column_unions = [&](uint64_t j)
    {
        if (!unions)
            return true;
        auto const t0 = std::chrono::steady_clock::now();
        union_col = unions->column(j);
        sk->union_seconds += seconds_since(t0);
        return union_col != nullptr;
    };
    if (nthreads == 1)
    {
        for (uint64_t j = 1; j < num_user_bins; ++j)
        {
            if (!column_unions(j))
                throw std::runtime_error("union estimates failed");
            column_split(j);
            for (uint64_t i = 1; i < num_technical_bins; ++i)
                cell(i, j);
            for (uint64_t i = 0; i < num_technical_bins; ++i)
            {
                colM[i] = matrix[i][j];
                colL[i] = ll[i][j];
            }
        }
    }
    else
    {
        std::barrier sync(static_cast<std::ptrdiff_t>(nthreads));
        bool failed = false;
        auto worker = [&](unsigned t)
        {
            for (uint64_t j = 1; j < num_user_bins; ++j)
            {
                if (t == 0)
                {
                    failed = !column_unions(j);
                    column_split(j);
                }
                sync.arrive_and_wait();
                if (failed)
                    return;
                for (uint64_t i = 1 + t; i < num_technical_bins; i += nthreads)
                    cell(i, j);
                sync.arrive_and_wait();
                for (uint64_t i = t; i < num_technical_bins; i += nthreads)
                {
                    colM[i] = matrix[i][j];
                    colL[i] = ll[i][j];
                }
                sync.arrive_and_wait();
            }
        };
        std::vector<std::thread> pool;
        for (unsigned t = 1; t < nthreads; ++t)
            pool.emplace_back(worker, t);
        worker(0);
        for (auto & th : pool)
            th.join();
        if (failed)
            throw std::runtime_error("union estimates failed");
    }

    // backtracking (:302-331) with backtrack_merged_bin / backtrack_split_bin
    MaxTracker tracker;
    // :258-287: with sketches the merged bin's size is the estimate of the
    // union of its user bins' sketches.
    auto merged = [&](uint64_t tj, uint64_t next_j, uint64_t bin_id, bool is_first_row)
    {
        uint64_t kmer_count = counts[d.positions[tj]];
        seqan::hibf::sketch::hyperloglog merged_sketch =
            sk ? (*sk->sketches)[d.positions[tj]] : seqan::hibf::sketch::hyperloglog{};
        Store lower{d.layout, d.kmer_counts, {d.positions[tj]}, d.fpr_correction, d.relaxed_fpr_correction, d.previous,
                    d.sketching};
        lower.previous.push_back(bin_id);
        tj -= !is_first_row;
        while (tj != next_j)
        {
            tj -= is_first_row;
            if (sk)
                merged_sketch.merge((*sk->sketches)[d.positions[tj]]);
            else
                kmer_count += counts[d.positions[tj]];
            lower.positions.push_back(d.positions[tj]);
            tj -= !is_first_row;
        }
        std::vector<uint64_t> const prev = lower.previous;
        uint64_t const lower_max = add_lower_level(lower, c);
        d.layout->max_bins.push_back({prev, lower_max});
        if (sk)
            kmer_count = static_cast<uint64_t>(merged_sketch.estimate());
        tracker.update(bin_id, static_cast<uint64_t>(kmer_count * rcorr));
    };
    auto split = [&](uint64_t tj, uint64_t nb, uint64_t bin_id)
    {
        d.layout->user_bins.push_back({d.previous, bin_id, nb, d.positions[tj]});
        uint64_t const card = counts[d.positions[tj]];
        uint64_t const corrected = static_cast<uint64_t>(card * corr[nb]);
        tracker.update(bin_id, divide_and_ceil(corrected, nb));
    };
    uint64_t trace_i = num_technical_bins - 1;
    uint64_t trace_j = num_user_bins - 1;
    uint64_t bin_id{};
    while (trace_j > 0u && trace_i > 0u)
    {
        uint64_t const next_i = trace[trace_i][trace_j].first;
        uint64_t const next_j = trace[trace_i][trace_j].second;
        uint64_t const nb = trace_i - next_i;
        if (nb == 1 && next_j != trace_j - 1u)
        {
            merged(trace_j, next_j, bin_id, false);
            trace_i = next_i;
            trace_j = next_j;
        }
        else
        {
            split(trace_j, nb, bin_id);
            trace_i = trace[trace_i][trace_j].first;
            --trace_j;
        }
        bin_id += nb;
    }
    if (trace_i == 0u && trace_j > 0u)
        merged(trace_j, 0, bin_id, true);
    else if (trace_j == 0u)
        split(trace_j, trace_i + 1, bin_id);
    return tracker.max_id;
}

} // namespace

uint64_t simple_binning_layout(std::vector<uint64_t> const & kmer_counts,
                               std::vector<uint64_t> const & positions,
                               std::vector<double> const & fpr_correction,
                               uint64_t num_technical_bins,
                               Layout & layout)
{
    if (positions.size() > num_technical_bins)
        throw std::runtime_error("You cannot have less technical bins than user bins for this simple binning "
                                 "algorithm. Please see the hierarchical_binning algorithm or increase the number "
                                 "of technical bins.");
    Store store{&layout, &kmer_counts, positions, &fpr_correction, 1.0, {}, nullptr};
    return simple_binning(store, num_technical_bins);
}

bool HostUnionEstimates::start(std::vector<uint64_t> const & positions)
{
    positions_ = positions;
    return true;
}

bool HostUnionEstimates::prefix(uint64_t * prefix)
{
    if (positions_.empty())
        return true;
    seqan::hibf::sketch::hyperloglog running = sketches_[positions_[0]];
    for (uint64_t j = 1; j < positions_.size(); ++j)
        prefix[j] = static_cast<uint64_t>(running.merge_and_estimate(sketches_[positions_[j]]));
    return true;
}

uint64_t const * HostUnionEstimates::column(uint64_t j)
{
    column_.resize(j);
    seqan::hibf::sketch::hyperloglog running = sketches_[positions_[j]];
    for (uint64_t k = j; k-- > 0;)
        column_[k] = static_cast<uint64_t>(running.merge_and_estimate(sketches_[positions_[k]]));
    return column_.data();
}

Layout compute_layout(Config const & c,
                      std::vector<uint64_t> const & kmer_counts,
                      Sketching * sketching,
                      std::vector<uint64_t> const * positions)
{
    Layout layout;
    std::vector<double> const corr = fpr_correction(c.maximum_fpr, c.number_of_hash_functions, c.tmax);
    if (c.disable_estimate_union || (sketching && !sketching->sketches))
        sketching = nullptr;
    Store store{&layout, &kmer_counts, {}, &corr,
                relaxed_fpr_correction(c.maximum_fpr, c.relaxed_fpr, c.number_of_hash_functions), {}, sketching};
    if (positions)
        store.positions = *positions;
    else
    {
        store.positions.resize(c.number_of_user_bins);
        for (uint64_t i = 0; i < c.number_of_user_bins; ++i)
            store.positions[i] = i;
    }
    layout.top_level_max_bin_id = hierarchical_binning(store, c);
    // compute_layout.cpp:61-67: max bins by depth (stable here; the order of
    // equal depths does not change the built filter's answers).
    std::stable_sort(layout.max_bins.begin(), layout.max_bins.end(),
                     [](Layout::MaxBin const & a, Layout::MaxBin const & b)
                     { return a.previous_TB_indices.size() < b.previous_TB_indices.size(); });
    return layout;
}

std::string Layout::to_text() const
{
    std::ostringstream os;
    os << "#TOP_LEVEL_IBF fullest_technical_bin_idx:" << top_level_max_bin_id << '\n';
    for (auto const & m : max_bins)
    {
        os << "#LOWER_LEVEL_IBF_";
        for (size_t i = 0; i < m.previous_TB_indices.size(); ++i)
            os << (i ? ";" : "") << m.previous_TB_indices[i];
        os << " fullest_technical_bin_idx:" << m.id << '\n';
    }
    os << "#USER_BIN_IDX\tTECHNICAL_BIN_INDICES\tNUMBER_OF_TECHNICAL_BINS\n";
    for (auto const & u : user_bins)
    {
        os << u.idx << '\t';
        for (auto b : u.previous_TB_indices)
            os << b << ';';
        os << u.storage_TB_id << '\t';
        for (size_t i = 0; i < u.previous_TB_indices.size(); ++i)
            os << "1;";
        os << u.number_of_technical_bins << '\n';
    }
    return os.str();
}

namespace
{
std::vector<uint64_t> parse_list(std::string const & s, char sep)
{
    std::vector<uint64_t> out;
    std::string tok;
    std::istringstream is(s);
    while (std::getline(is, tok, sep))
        if (!tok.empty())
            out.push_back(std::stoull(tok));
    return out;
}
} // namespace

bool Layout::from_text(std::string const & text, Layout & out)
{
    out = Layout{};
    std::istringstream is(text);
    std::string line;
    // Skip a leading @HIBF_CONFIG block if present (layout files carry the config first).
    while (std::getline(is, line))
        if (!line.empty() && line[0] != '@')
            break;
    std::string const top = "#TOP_LEVEL_IBF fullest_technical_bin_idx:";
    if (line.rfind(top, 0) != 0)
        return false;
    try
    {
        out.top_level_max_bin_id = std::stoull(line.substr(top.size()));
        std::string const lower = "#LOWER_LEVEL_IBF_";
        std::string const cols = "#USER_BIN_IDX\tTECHNICAL_BIN_INDICES\tNUMBER_OF_TECHNICAL_BINS";
        while (std::getline(is, line) && line != cols)
        {
            if (line.rfind(lower, 0) != 0)
                return false;
            size_t const sp = line.find(' ');
            size_t const colon = line.find(':', sp);
            if (sp == std::string::npos || colon == std::string::npos)
                return false;
            out.max_bins.push_back({parse_list(line.substr(lower.size(), sp - lower.size()), ';'),
                                    std::stoull(line.substr(colon + 1))});
        }
        if (line != cols)
            return false;
        while (std::getline(is, line))
        {
            if (line.empty() || line[0] == '#')
                break;
            std::istringstream ls(line);
            std::string f0, f1, f2;
            if (!std::getline(ls, f0, '\t') || !std::getline(ls, f1, '\t') || !std::getline(ls, f2, '\t'))
                return false;
            UserBin u;
            u.idx = std::stoull(f0);
            std::vector<uint64_t> tbs = parse_list(f1, ';');
            std::vector<uint64_t> nbs = parse_list(f2, ';');
            if (tbs.empty() || nbs.empty())
                return false;
            u.storage_TB_id = tbs.back();
            tbs.pop_back();
            u.previous_TB_indices = tbs;
            u.number_of_technical_bins = nbs.back();
            out.user_bins.push_back(std::move(u));
        }
    }
    catch (...)
    {
        return false;
    }
    return true;
}

} // namespace hibf_amd::host
