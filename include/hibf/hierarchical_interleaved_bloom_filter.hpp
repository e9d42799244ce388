// seqan::hibf::hierarchical_interleaved_bloom_filter for the MI355X-native
// query path.  Same names and argument meaning as the reference
// (hierarchical_interleaved_bloom_filter.hpp:32-580): construction from a
// config (layout on the host, bit matrices by the device bulk emplace),
// membership_agent().membership_for(values, threshold) in the reference's
// unsorted depth-first order, sort_results(), and
// counting_agent<T>().bulk_count(values, threshold).  Batched forms
// (membership_for_batch / bulk_count_batch) send many queries in one launch.
#pragma once

#include <algorithm>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <filesystem>
#include <istream>
#include <iterator>
#include <memory>
#include <ranges>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <hibf/config.hpp>
#include <hibf/detail/device_buffer.hpp>
#include <hibf/interleaved_bloom_filter.hpp>
#include <hibf/layout/layout.hpp>
#include <hibf/misc/timer.hpp>
#include <hibf/misc/counting_vector.hpp>
#include <hibf_gpu.h>

namespace seqan::hibf
{

namespace bin_kind
{
static constexpr uint64_t merged{HIBF_GPU_BIN_MERGED};
static constexpr uint64_t deleted{HIBF_GPU_BIN_DELETED};
} // namespace bin_kind


namespace detail
{
struct hibf_handle
{
    hibf_gpu_hibf h{nullptr};
    explicit hibf_handle(hibf_gpu_hibf x) : h{x} {}
    ~hibf_handle() { hibf_gpu_hibf_destroy(h); }
};

// Device workspace of one agent, created on first use; a copied agent gets
// its own workspace, so copies can be used from different threads as the
// reference allows.
struct agent_handle
{
    agent_handle() = default;
    agent_handle(agent_handle const &) noexcept {}
    agent_handle & operator=(agent_handle const &) noexcept { return *this; }
    ~agent_handle() { hibf_gpu_agent_destroy(a_); }
    hibf_gpu_agent get()
    {
        if (!a_)
            check(hibf_gpu_agent_create(0, &a_));
        return a_;
    }

private:
    hibf_gpu_agent a_{nullptr};
};

// The batched agents take one threshold per query; anything else would send
// the device kernels past the end of the thresholds buffer.
template <typename queries_t, typename thresholds_t>
void check_batch_thresholds(queries_t const & queries, thresholds_t const & thresholds)
{
    auto const nq = static_cast<size_t>(std::ranges::distance(queries));
    if (thresholds.size() != nq)
        throw std::invalid_argument{"one threshold per query is required (" + std::to_string(thresholds.size()) +
                                    " thresholds for " + std::to_string(nq) + " queries)"};
}

// Uploads n elements of T (one zero element when n == 0, so that the device
// buffer exists and no empty vector's data() is read).
template <typename T>
void upload_or_zero(device_buffer & d, std::vector<T> const & h)
{
    T const zero{};
    d.upload(h.empty() ? &zero : h.data(), std::max<size_t>(h.size(), 1) * sizeof(T));
}

// Flattens a batch of queries into CSR form on the device.
struct query_batch
{
    device_buffer values, offsets;
    size_t nq{0};

    template <typename queries_t>
    void upload(queries_t const & queries)
    {
        std::vector<uint64_t> v, o{0};
        for (auto const & q : queries)
        {
            for (auto && x : q)
                v.push_back(static_cast<uint64_t>(x));
            o.push_back(v.size());
        }
        nq = o.size() - 1;
        if (v.empty())
            v.push_back(0);
        values.upload(v.data(), v.size() * 8u);
        offsets.upload(o.data(), o.size() * 8u);
    }
};
} // namespace detail

class hierarchical_interleaved_bloom_filter
{
public:
    class membership_agent_type;
    template <std::integral value_t>
    class counting_agent_type;

    // hierarchical_interleaved_bloom_filter.hpp:292-302
    struct previous_ibf_id_pair
    {
        size_t ibf_idx{};
        size_t bin_idx{};

        friend constexpr auto operator<=>(previous_ibf_id_pair const &, previous_ibf_id_pair const &) = default;
    };

    size_t number_of_user_bins{};
    // ibf_vector[i] is a view of IBF i in the device arena (no copy).
    std::vector<interleaved_bloom_filter> ibf_vector;
    std::vector<std::vector<uint64_t>> next_ibf_id;
    // Parent (IBF, bin) of each IBF, {0, 0} for the root (hierarchical_interleaved_bloom_filter.hpp:220-229).
    std::vector<previous_ibf_id_pair> prev_ibf_id;
    std::vector<std::vector<uint64_t>> ibf_bin_to_user_bin_id;

    static constexpr uint32_t version{1};

    // Build phases (hierarchical_interleaved_bloom_filter.hpp:283-290).  The device build reports sketches,
    // layout, allocation and fill; reading the user bins is timed here.  Union estimation and
    // rearrangement are not performed, merging k-mers is part of the device fill: those stay zero.
    concurrent_timer layout_compute_sketches_timer{};
    concurrent_timer layout_union_estimation_timer{};
    concurrent_timer layout_rearrangement_timer{};
    concurrent_timer layout_dp_algorithm_timer{};
    concurrent_timer index_allocation_timer{};
    concurrent_timer user_bin_io_timer{};
    concurrent_timer merge_kmers_timer{};
    concurrent_timer fill_ibf_timer{};

    hierarchical_interleaved_bloom_filter() = default;

    // Copies share the device HIBF (its arena and tables stay on the device once); the copy's
    // ibf_vector views the same arena.
    hierarchical_interleaved_bloom_filter(hierarchical_interleaved_bloom_filter const & o) { *this = o; }
    hierarchical_interleaved_bloom_filter & operator=(hierarchical_interleaved_bloom_filter const & o)
    {
        if (this == &o)
            return *this;
        if (o.handle_)
            share(o.handle_);
        else
            *this = hierarchical_interleaved_bloom_filter{};
        layout_compute_sketches_timer = o.layout_compute_sketches_timer;
        layout_union_estimation_timer = o.layout_union_estimation_timer;
        layout_rearrangement_timer = o.layout_rearrangement_timer;
        layout_dp_algorithm_timer = o.layout_dp_algorithm_timer;
        index_allocation_timer = o.index_allocation_timer;
        user_bin_io_timer = o.user_bin_io_timer;
        merge_kmers_timer = o.merge_kmers_timer;
        fill_ibf_timer = o.fill_ibf_timer;
        return *this;
    }
    hierarchical_interleaved_bloom_filter(hierarchical_interleaved_bloom_filter &&) noexcept = default;
    hierarchical_interleaved_bloom_filter & operator=(hierarchical_interleaved_bloom_filter &&) noexcept = default;
    ~hierarchical_interleaved_bloom_filter() = default;

    // src/hierarchical_interleaved_bloom_filter.cpp:211-244
    hierarchical_interleaved_bloom_filter(config & configuration) { build(configuration, nullptr); }

    // src/hierarchical_interleaved_bloom_filter.cpp:246-253
    hierarchical_interleaved_bloom_filter(config & configuration, layout::layout const & l)
    {
        std::ostringstream text;
        l.write_to(text);
        std::string const s = text.str();
        build(configuration, &s);
    }

    membership_agent_type membership_agent() const;
    template <std::integral value_t = uint16_t>
    counting_agent_type<value_t> counting_agent() const;

    hibf_gpu_hibf handle() const noexcept { return handle_ ? handle_->h : nullptr; }

    // Archive in the byte layout a cereal::BinaryOutputArchive writes for the
    // reference's object (serialize, hierarchical_interleaved_bloom_filter.hpp:267-277);
    // `append` adds it to the end of an existing file.
    void store(std::filesystem::path const & path, bool const append = false) const
    {
        detail::check(hibf_gpu_hibf_save(handle(), path.c_str(), append ? 1 : 0));
    }

    // Reads such an archive starting at byte `offset`; returns the offset after it.
    uint64_t load(std::filesystem::path const & path, uint64_t const offset = 0u)
    {
        hibf_gpu_hibf h = nullptr;
        uint64_t end = 0;
        detail::check(hibf_gpu_hibf_load(path.c_str(), offset, &h, &end));
        adopt(h);
        return end;
    }

private:
    void build(config & configuration, std::string const * layout_text)
    {
        configuration.validate_and_set_defaults();
        std::vector<uint64_t> values, offsets{0};
        std::vector<uint64_t> scratch;
        user_bin_io_timer.start();
        for (size_t i = 0; i < configuration.number_of_user_bins; ++i)
        {
            scratch.clear();
            configuration.input_fn(i, insert_iterator{scratch});
            values.insert(values.end(), scratch.begin(), scratch.end());
            offsets.push_back(values.size());
        }
        user_bin_io_timer.stop();
        hibf_gpu_config const c = configuration.to_c();
        hibf_gpu_hibf h = nullptr;
        detail::check(hibf_gpu_hibf_build(&c, offsets.data(), values.empty() ? offsets.data() : values.data(),
                                          layout_text ? layout_text->c_str() : nullptr, &h));
        adopt(h);
        double phase[4]{};
        detail::check(hibf_gpu_hibf_build_times(h, phase));
        if (!layout_text)
        {
            double parts[2]{};
            detail::check(hibf_gpu_hibf_layout_times(h, parts));
            layout_compute_sketches_timer.add_seconds(phase[0]);
            layout_union_estimation_timer.add_seconds(parts[0]);
            layout_rearrangement_timer.add_seconds(parts[1]);
        }
        layout_dp_algorithm_timer.add_seconds(phase[1]);
        index_allocation_timer.add_seconds(phase[2]);
        fill_ibf_timer.add_seconds(phase[3]);
    }

    // Takes ownership of a device HIBF and mirrors its routing tables.
    void adopt(hibf_gpu_hibf h) { share(std::make_shared<detail::hibf_handle>(h)); }

    // Mirrors the routing tables of a (shared) device HIBF and views its IBFs.
    void share(std::shared_ptr<detail::hibf_handle> handle)
    {
        handle_ = std::move(handle);
        hibf_gpu_hibf const h = handle_->h;
        uint64_t n_ibfs = 0, n_tbs = 0, n_ub = 0;
        detail::check(hibf_gpu_hibf_info(h, &n_ibfs, &n_tbs, nullptr, &n_ub, nullptr));
        number_of_user_bins = n_ub;
        std::vector<hibf_gpu_ibf_geometry> geoms(n_ibfs);
        std::vector<uint64_t> wo(n_ibfs), to(n_ibfs + 1), nx(n_tbs), ub(n_tbs);
        detail::check(hibf_gpu_hibf_export_tables(h, geoms.data(), wo.data(), to.data(), nx.data(), ub.data()));
        next_ibf_id.assign(n_ibfs, {});
        ibf_bin_to_user_bin_id.assign(n_ibfs, {});
        prev_ibf_id.assign(n_ibfs, previous_ibf_id_pair{});
        ibf_vector.clear();
        ibf_vector.reserve(n_ibfs);
        for (uint64_t i = 0; i < n_ibfs; ++i)
        {
            next_ibf_id[i].assign(nx.begin() + to[i], nx.begin() + to[i + 1]);
            ibf_bin_to_user_bin_id[i].assign(ub.begin() + to[i], ub.begin() + to[i + 1]);
            for (uint64_t b = 0; b < to[i + 1] - to[i]; ++b)
                if (ub[to[i] + b] == HIBF_GPU_BIN_MERGED)
                    prev_ibf_id[nx[to[i] + b]] = previous_ibf_id_pair{i, b};
            hibf_gpu_ibf view = nullptr;
            detail::check(hibf_gpu_hibf_ibf_view(h, i, &view));
            ibf_vector.push_back(interleaved_bloom_filter::view_of(view, handle_));
        }
    }

    std::shared_ptr<detail::hibf_handle> handle_{};
};

// hierarchical_interleaved_bloom_filter.hpp:314-446
class hierarchical_interleaved_bloom_filter::membership_agent_type
{
public:
    membership_agent_type() = default;
    explicit membership_agent_type(hierarchical_interleaved_bloom_filter const & hibf) :
        hibf_ptr{std::addressof(hibf)}
    {}

    void sort_results() { std::ranges::sort(result_buffer); }

    template <std::ranges::forward_range value_range_t>
    [[nodiscard]] std::vector<uint64_t> const & membership_for(value_range_t && values,
                                                               uint16_t const threshold) & noexcept
    {
        std::vector<std::vector<uint64_t>> one{detail::to_u64(values)};
        run(one, std::vector<uint16_t>{threshold});
        result_buffer.assign(ids_.begin(), ids_.end());
        return result_buffer;
    }
    template <std::ranges::forward_range value_range_t>
    [[nodiscard]] std::vector<uint64_t> const & membership_for(value_range_t && values,
                                                               uint16_t const threshold) && noexcept = delete;

    // Batched form: one launch for all queries; result i is query i's ids in
    // the reference's order.
    template <typename queries_t>
    std::vector<std::vector<uint64_t>> membership_for_batch(queries_t const & queries,
                                                            std::vector<uint16_t> const & thresholds)
    {
        detail::check_batch_thresholds(queries, thresholds);
        run(queries, thresholds);
        std::vector<std::vector<uint64_t>> out(batch_.nq);
        for (size_t q = 0; q < batch_.nq; ++q)
            out[q].assign(ids_.begin() + hoff_[q], ids_.begin() + hoff_[q + 1]);
        return out;
    }

private:
    template <typename queries_t>
    void run(queries_t const & queries, std::vector<uint16_t> const & thresholds) noexcept
    {
        try
        {
            batch_.upload(queries);
            detail::upload_or_zero(d_thr_, thresholds);
            d_hoff_.reserve((batch_.nq + 1) * 8u);
            uint64_t total = 0;
            uint64_t cap = std::max<uint64_t>(cap_, 1024u);
            for (;;)
            {
                d_ids_.reserve(cap * 8u);
                int const rc = hibf_gpu_hibf_membership_for(hibf_ptr->handle(), agent_.get(), batch_.values.as<uint64_t>(),
                                                            batch_.offsets.as<uint64_t>(), batch_.nq,
                                                            d_thr_.as<uint16_t>(), d_ids_.as<uint64_t>(), cap,
                                                            d_hoff_.as<uint64_t>(), &total, nullptr);
                if (rc == HIBF_GPU_ERR_CAPACITY)
                {
                    cap = total;
                    continue;
                }
                detail::check(rc);
                break;
            }
            cap_ = cap;
            ids_.resize(total);
            hoff_.resize(batch_.nq + 1);
            detail::check(hibf_gpu_synchronize());
            d_ids_.download(ids_.data(), total * 8u);
            d_hoff_.download(hoff_.data(), hoff_.size() * 8u);
        }
        catch (...)
        {
            std::terminate();
        }
    }

    hierarchical_interleaved_bloom_filter const * hibf_ptr{nullptr};
    detail::agent_handle agent_;
    detail::query_batch batch_;
    detail::device_buffer d_thr_, d_ids_, d_hoff_;
    uint64_t cap_{0};
    std::vector<uint64_t> ids_, hoff_;
    std::vector<uint64_t> result_buffer;
};

// hierarchical_interleaved_bloom_filter.hpp:454-573
template <std::integral value_t>
class hierarchical_interleaved_bloom_filter::counting_agent_type
{
public:
    counting_agent_type() = default;
    explicit counting_agent_type(hierarchical_interleaved_bloom_filter const & hibf) :
        hibf_ptr{std::addressof(hibf)},
        result_buffer(hibf.number_of_user_bins)
    {}

    template <std::ranges::forward_range value_range_t>
    [[nodiscard]] counting_vector<value_t> const & bulk_count(value_range_t && values,
                                                              size_t const threshold) & noexcept
    {
        std::vector<std::vector<uint64_t>> one{detail::to_u64(values)};
        auto const counts = bulk_count_batch(one, std::vector<uint64_t>{threshold});
        std::copy(counts.begin(), counts.end(), result_buffer.begin());
        return result_buffer;
    }
    template <std::ranges::forward_range value_range_t>
    [[nodiscard]] counting_vector<value_t> const & bulk_count(value_range_t && values,
                                                              size_t const threshold) && noexcept = delete;

    // Batched form: row q of the result (number_of_user_bins counters) is query q's counts.
    // std::invalid_argument unless there is one threshold per query.
    template <typename queries_t>
    std::vector<value_t> bulk_count_batch(queries_t const & queries, std::vector<uint64_t> const & thresholds)
    {
        detail::check_batch_thresholds(queries, thresholds);
        try
        {
            batch_.upload(queries);
            detail::upload_or_zero(d_thr_, thresholds);
            size_t const n = batch_.nq * hibf_ptr->number_of_user_bins;
            d_counts_.reserve(std::max<size_t>(n, 1) * sizeof(value_t));
            detail::check(hibf_gpu_hibf_bulk_count(hibf_ptr->handle(), agent_.get(), batch_.values.as<uint64_t>(),
                                                   batch_.offsets.as<uint64_t>(), batch_.nq, d_thr_.as<uint64_t>(),
                                                   sizeof(value_t), d_counts_.as<void>(), nullptr));
            detail::check(hibf_gpu_synchronize());
            std::vector<value_t> out(n);
            d_counts_.download(out.data(), n * sizeof(value_t));
            return out;
        }
        catch (...)
        {
            std::terminate();
        }
    }

private:
    hierarchical_interleaved_bloom_filter const * hibf_ptr{nullptr};
    detail::agent_handle agent_;
    detail::query_batch batch_;
    detail::device_buffer d_thr_, d_counts_;
    counting_vector<value_t> result_buffer;
};

inline hierarchical_interleaved_bloom_filter::membership_agent_type
hierarchical_interleaved_bloom_filter::membership_agent() const
{
    return membership_agent_type{*this};
}

template <std::integral value_t>
inline hierarchical_interleaved_bloom_filter::counting_agent_type<value_t>
hierarchical_interleaved_bloom_filter::counting_agent() const
{
    return counting_agent_type<value_t>{*this};
}

} // namespace seqan::hibf
