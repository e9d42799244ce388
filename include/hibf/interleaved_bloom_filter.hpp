// seqan::hibf::interleaved_bloom_filter for the MI355X-native query path.
//
// Same names, argument meaning and error behaviour as the reference
// (interleaved_bloom_filter.hpp:51-801): the bit matrix lives in HBM (one
// hibf_gpu_ibf handle), emplace/clear/bulk_contains/bulk_count/membership_for
// run as HIP kernels through the C-ABI of hibf_gpu.h.  Agents own their
// result buffers and return const references that the next call
// invalidates, as in the reference.  Query functions are noexcept; a device
// failure inside them terminates, as a broken invariant would.
#pragma once

#include <algorithm>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <filesystem>
#include <iterator>
#include <memory>
#include <ranges>
#include <stdexcept>
#include <unordered_set>
#include <vector>

#include <hibf/config.hpp>
#include <hibf/detail/device_buffer.hpp>
#include <hibf/misc/bit_vector.hpp>
#include <hibf/misc/counting_vector.hpp>
#include <hibf/misc/insert_iterator.hpp>
#include <hibf_gpu.h>

namespace seqan::hibf
{

// interleaved_bloom_filter.hpp:51-81
struct bin_count
{
    size_t value;
};
struct bin_size
{
    size_t value;
};
struct hash_function_count
{
    size_t value;
};
struct bin_index
{
    size_t value;
};

namespace detail
{
template <typename range_t>
std::vector<uint64_t> to_u64(range_t && values)
{
    std::vector<uint64_t> out;
    if constexpr (std::ranges::sized_range<range_t>)
        out.reserve(std::ranges::size(values));
    for (auto && v : values)
        out.push_back(static_cast<uint64_t>(v));
    return out;
}

struct ibf_handle
{
    hibf_gpu_ibf h{nullptr};
    std::shared_ptr<void const> owner{}; // what a view's words belong to (an HIBF), kept alive
    explicit ibf_handle(hibf_gpu_ibf x, std::shared_ptr<void const> o = {}) : h{x}, owner{std::move(o)} {}
    ~ibf_handle() { hibf_gpu_ibf_destroy(h); }
};
} // namespace detail

class interleaved_bloom_filter
{
public:
    class containment_agent_type;
    template <std::integral value_t>
    class counting_agent_type;
    class membership_agent_type;

    interleaved_bloom_filter() = default;

    interleaved_bloom_filter(seqan::hibf::bin_count const bins_,
                             seqan::hibf::bin_size const size,
                             seqan::hibf::hash_function_count const funs = seqan::hibf::hash_function_count{2u},
                             bool const track_occupancy_ = false) :
        track_occupancy{track_occupancy_}
    {
        hibf_gpu_ibf h = nullptr;
        detail::check(hibf_gpu_ibf_create(bins_.value, size.value, static_cast<uint32_t>(funs.value), &h));
        handle_ = std::make_shared<detail::ibf_handle>(h);
        detail::check(hibf_gpu_ibf_get_geometry(h, &geometry_));
        occupancy.assign(geometry_.technical_bins, 0u);
    }

    // src/interleaved_bloom_filter.cpp:81-117: bin size from the largest user
    // bin (exact distinct count unless max_bin_elements is given), then every
    // user bin is inserted into its technical bin.
    interleaved_bloom_filter(config & configuration, size_t const max_bin_elements = 0u)
    {
        configuration.validate_and_set_defaults();
        std::vector<std::vector<uint64_t>> bins(configuration.number_of_user_bins);
        size_t max_size = max_bin_elements;
        for (size_t i = 0; i < bins.size(); ++i)
        {
            configuration.input_fn(i, insert_iterator{bins[i]});
            if (max_bin_elements == 0u)
            {
                std::unordered_set<uint64_t> distinct(bins[i].begin(), bins[i].end());
                max_size = std::max(max_size, distinct.size());
            }
        }
        uint64_t const bits = hibf_gpu_bin_size_in_bits(configuration.maximum_fpr,
                                                        configuration.number_of_hash_functions, max_size);
        *this = interleaved_bloom_filter{seqan::hibf::bin_count{configuration.number_of_user_bins},
                                         seqan::hibf::bin_size{bits},
                                         seqan::hibf::hash_function_count{configuration.number_of_hash_functions},
                                         configuration.track_occupancy};
        for (size_t i = 0; i < bins.size(); ++i)
            emplace_range(bins[i], seqan::hibf::bin_index{i});
    }

    // Copies share nothing: the device matrix is duplicated.
    interleaved_bloom_filter(interleaved_bloom_filter const & o) :
        track_occupancy{o.track_occupancy},
        occupancy{o.occupancy},
        geometry_{o.geometry_}
    {
        if (o.handle_)
        {
            hibf_gpu_ibf h = nullptr;
            detail::check(hibf_gpu_ibf_create(o.geometry_.bins, o.geometry_.bin_size,
                                              static_cast<uint32_t>(o.geometry_.hash_funs), &h));
            handle_ = std::make_shared<detail::ibf_handle>(h);
            uint64_t * src = nullptr;
            uint64_t * dst = nullptr;
            detail::check(hibf_gpu_ibf_device_words(o.handle_->h, &src));
            detail::check(hibf_gpu_ibf_device_words(h, &dst));
            detail::check(hibf_gpu_memcpy(dst, src, o.geometry_.n_words * 8u, 2));
        }
    }
    interleaved_bloom_filter & operator=(interleaved_bloom_filter const & o)
    {
        if (this != &o)
        {
            interleaved_bloom_filter tmp{o};
            *this = std::move(tmp);
        }
        return *this;
    }
    interleaved_bloom_filter(interleaved_bloom_filter &&) noexcept = default;
    interleaved_bloom_filter & operator=(interleaved_bloom_filter &&) noexcept = default;
    ~interleaved_bloom_filter() = default;

    // interleaved_bloom_filter.hpp:259 / src/interleaved_bloom_filter.cpp:119-140
    void emplace(size_t const value, bin_index const bin) noexcept
    {
        uint64_t const v = value;
        emplace_range(std::vector<uint64_t>{v}, bin);
    }

    // Bulk form: every value of the range goes to `bin` in one device launch.
    template <std::ranges::range range_t>
    void emplace_range(range_t && values, bin_index const bin) noexcept
    {
        std::vector<uint64_t> const v = detail::to_u64(values);
        if (v.empty())
            return;
        detail::device_buffer d;
        d.upload(v.data(), v.size() * 8u);
        int rc = HIBF_GPU_OK;
        if (track_occupancy)
        {
            // The handle counts the values that set a new bit, in range order, as that many emplace calls
            // of the reference do; its counters start from this object's public occupancy.
            std::vector<uint64_t> occ(occupancy.begin(), occupancy.end());
            rc = hibf_gpu_ibf_set_occupancy(handle_->h, occ.data(), 1);
            if (rc == HIBF_GPU_OK)
                rc = hibf_gpu_ibf_emplace(handle_->h, d.as<uint64_t>(), nullptr, bin.value, v.size(), nullptr);
            if (rc == HIBF_GPU_OK)
                rc = hibf_gpu_ibf_occupancy(handle_->h, occ.data(), nullptr);
            if (rc == HIBF_GPU_OK)
                occupancy.assign(occ.begin(), occ.end());
        }
        else
        {
            rc = hibf_gpu_ibf_emplace(handle_->h, d.as<uint64_t>(), nullptr, bin.value, v.size(), nullptr);
        }
        if (rc == HIBF_GPU_OK)
            rc = hibf_gpu_synchronize();
        if (rc != HIBF_GPU_OK)
            std::terminate();
    }

    // interleaved_bloom_filter.hpp:270, 283-297
    void clear(bin_index const bin) noexcept
    {
        uint64_t const b = bin.value;
        if (hibf_gpu_ibf_clear(handle_->h, &b, 1, nullptr) != HIBF_GPU_OK)
            std::terminate();
        occupancy[bin.value] = 0u;
    }
    template <typename rng_t>
        requires std::same_as<std::ranges::range_value_t<rng_t>, bin_index>
    void clear(rng_t && bin_range) noexcept
    {
        std::vector<uint64_t> bins;
        for (auto && b : bin_range)
            bins.push_back(b.value);
        if (hibf_gpu_ibf_clear(handle_->h, bins.data(), bins.size(), nullptr) != HIBF_GPU_OK)
            std::terminate();
        for (uint64_t b : bins)
            occupancy[b] = 0u;
    }

    // src/interleaved_bloom_filter.cpp:150-160
    // src/interleaved_bloom_filter.cpp:150-160: only bin_count changes.
    bool try_increase_bin_number_to(seqan::hibf::bin_count const new_bin_count) noexcept
    {
        int increased = 0;
        if (!handle_ || hibf_gpu_ibf_try_increase_bin_number_to(handle_->h, new_bin_count.value, &increased) != HIBF_GPU_OK)
            return false;
        if (increased)
            geometry_.bins = new_bin_count.value;
        return increased != 0;
    }

    // src/interleaved_bloom_filter.cpp:163-201: the rows are re-strided to the
    // new technical-bin count on the device.
    void increase_bin_number_to(seqan::hibf::bin_count const new_bin_count)
    {
        if (new_bin_count.value < geometry_.bins)
            throw std::invalid_argument{"The number of new bins must be >= the current number of bins."};
        detail::check(hibf_gpu_ibf_increase_bin_number_to(handle_->h, new_bin_count.value));
        detail::check(hibf_gpu_ibf_get_geometry(handle_->h, &geometry_));
        occupancy.resize(geometry_.technical_bins, 0u);
    }

    containment_agent_type containment_agent() const;
    template <std::integral value_t = uint16_t>
    counting_agent_type<value_t> counting_agent() const;
    membership_agent_type membership_agent() const;

    size_t hash_function_count() const noexcept { return geometry_.hash_funs; }
    size_t bin_count() const noexcept { return geometry_.bins; }
    size_t bin_size() const noexcept { return geometry_.bin_size; }
    size_t bit_size() const noexcept { return geometry_.technical_bins * geometry_.bin_size; }

    // The reference's data() is a host pointer to the bit_vector words; here
    // the words live in HBM, so there is no data(): device_data() is their
    // device address and words() a host copy.
    uint64_t * device_data() const
    {
        uint64_t * d = nullptr;
        detail::check(hibf_gpu_ibf_device_words(handle_->h, &d));
        return d;
    }

    // Host copy of the bit matrix (bit_vector words, row-major).
    std::vector<uint64_t> words() const
    {
        std::vector<uint64_t> w(geometry_.n_words);
        detail::check(hibf_gpu_ibf_download(handle_->h, w.data(), w.size()));
        return w;
    }

    bool operator==(interleaved_bloom_filter const & o) const
    {
        return geometry_.bins == o.geometry_.bins && geometry_.bin_size == o.geometry_.bin_size
            && geometry_.hash_funs == o.geometry_.hash_funs && words() == o.words();
    }

    hibf_gpu_ibf handle() const noexcept { return handle_ ? handle_->h : nullptr; }
    hibf_gpu_ibf_geometry const & geometry() const noexcept { return geometry_; }

    // An IBF over words it does not own: hierarchical_interleaved_bloom_filter::ibf_vector[i]
    // (hibf_gpu_hibf_ibf_view).  `owner` is kept alive as long as the view; copying the view
    // duplicates the matrix, as for any IBF.  Its bin count cannot change.
    static interleaved_bloom_filter view_of(hibf_gpu_ibf view, std::shared_ptr<void const> owner)
    {
        interleaved_bloom_filter ibf;
        ibf.handle_ = std::make_shared<detail::ibf_handle>(view, std::move(owner));
        detail::check(hibf_gpu_ibf_get_geometry(view, &ibf.geometry_));
        ibf.occupancy.assign(ibf.geometry_.technical_bins, 0u);
        return ibf;
    }

    // Archive in the byte layout a cereal::BinaryOutputArchive writes for the
    // reference's object (serialize, interleaved_bloom_filter.hpp:471-485);
    // `append` adds it to the end of an existing file.
    void store(std::filesystem::path const & path, bool const append = false) const
    {
        std::vector<uint64_t> const occ(occupancy.begin(), occupancy.end());
        detail::check(hibf_gpu_ibf_set_occupancy(handle_->h, occ.data(), track_occupancy ? 1 : 0));
        detail::check(hibf_gpu_ibf_save(handle_->h, path.c_str(), append ? 1 : 0));
    }

    // Reads such an archive starting at byte `offset`; returns the offset after it.
    uint64_t load(std::filesystem::path const & path, uint64_t const offset = 0u)
    {
        hibf_gpu_ibf h = nullptr;
        uint64_t end = 0;
        detail::check(hibf_gpu_ibf_load(path.c_str(), offset, &h, &end));
        handle_ = std::make_shared<detail::ibf_handle>(h);
        detail::check(hibf_gpu_ibf_get_geometry(h, &geometry_));
        std::vector<uint64_t> occ(geometry_.technical_bins);
        int track = 0;
        detail::check(hibf_gpu_ibf_occupancy(h, occ.data(), &track));
        occupancy.assign(occ.begin(), occ.end());
        track_occupancy = track != 0;
        return end;
    }

    bool track_occupancy{false};
    std::vector<size_t> occupancy{};

private:
    std::shared_ptr<detail::ibf_handle> handle_{};
    hibf_gpu_ibf_geometry geometry_{};
};

// interleaved_bloom_filter.hpp:499-557
class interleaved_bloom_filter::containment_agent_type
{
public:
    containment_agent_type() = default;
    explicit containment_agent_type(interleaved_bloom_filter const & ibf) :
        ibf_ptr{std::addressof(ibf)},
        result_buffer(ibf.bin_count())
    {}

    [[nodiscard]] bit_vector const & bulk_contains(size_t const value) & noexcept
    {
        uint64_t const v = value;
        size_t const words = ibf_ptr->geometry().bin_words;
        d_in.upload(&v, 8u);
        d_out.reserve(words * 8u);
        if (hibf_gpu_ibf_bulk_contains(ibf_ptr->handle(), d_in.as<uint64_t>(), 1, d_out.as<uint64_t>(), nullptr)
                != HIBF_GPU_OK
            || hibf_gpu_memcpy(result_buffer.data(), d_out.as<uint64_t>(), words * 8u, 1) != HIBF_GPU_OK)
            std::terminate();
        return result_buffer;
    }
    [[nodiscard]] bit_vector const & bulk_contains(size_t const value) && noexcept = delete;

private:
    interleaved_bloom_filter const * ibf_ptr{nullptr};
    bit_vector result_buffer;
    detail::device_buffer d_in, d_out;
};

// interleaved_bloom_filter.hpp:573-671 (counters of sizeof(value_t) bytes;
// signed types share the unsigned wrap-around bits)
template <std::integral value_t>
class interleaved_bloom_filter::counting_agent_type
{
public:
    counting_agent_type() = default;
    explicit counting_agent_type(interleaved_bloom_filter const & ibf) :
        ibf_ptr{std::addressof(ibf)},
        result_buffer(ibf.bin_count())
    {}

    template <std::ranges::range value_range_t>
    [[nodiscard]] counting_vector<value_t> const & bulk_count(value_range_t && values) & noexcept
    {
        static_assert(std::unsigned_integral<std::ranges::range_value_t<value_range_t>>,
                      "An individual value must be an unsigned integral.");
        std::vector<uint64_t> const v = detail::to_u64(values);
        uint64_t const offsets[2] = {0u, v.size()};
        d_values.upload(v.empty() ? offsets : static_cast<void const *>(v.data()), std::max<size_t>(v.size(), 1) * 8u);
        d_offsets.upload(offsets, 16u);
        d_counts.reserve(std::max<size_t>(result_buffer.size(), 1) * sizeof(value_t));
        if (hibf_gpu_ibf_bulk_count(ibf_ptr->handle(), d_values.as<uint64_t>(), d_offsets.as<uint64_t>(), 1,
                                    sizeof(value_t), d_counts.as<void>(), nullptr)
                != HIBF_GPU_OK
            || hibf_gpu_memcpy(result_buffer.data(), d_counts.as<void>(), result_buffer.size() * sizeof(value_t), 1)
                   != HIBF_GPU_OK)
            std::terminate();
        return result_buffer;
    }
    template <std::ranges::range value_range_t>
    [[nodiscard]] counting_vector<value_t> const & bulk_count(value_range_t && values) && noexcept = delete;

private:
    interleaved_bloom_filter const * ibf_ptr{nullptr};
    counting_vector<value_t> result_buffer;
    detail::device_buffer d_values, d_offsets, d_counts;
};

// interleaved_bloom_filter.hpp:683-794
class interleaved_bloom_filter::membership_agent_type
{
public:
    membership_agent_type() = default;
    explicit membership_agent_type(interleaved_bloom_filter const & ibf) : counting_agent{ibf}
    {
        result_buffer.reserve(ibf.bin_count());
    }

    void sort_results() {}

    template <std::ranges::forward_range value_range_t>
    [[nodiscard]] std::vector<uint64_t> const & membership_for(value_range_t && values,
                                                               uint16_t const threshold) & noexcept
    {
        result_buffer.clear();
        auto const & counts = counting_agent.bulk_count(std::forward<value_range_t>(values));
        for (size_t bin = 0; bin < counts.size(); ++bin)
            if (counts[bin] >= threshold)
                result_buffer.push_back(bin);
        return result_buffer;
    }
    template <std::ranges::forward_range value_range_t>
    [[nodiscard]] std::vector<uint64_t> const & membership_for(value_range_t && values,
                                                               uint16_t const threshold) && noexcept = delete;

private:
    counting_agent_type<uint16_t> counting_agent;
    std::vector<uint64_t> result_buffer;
};

inline interleaved_bloom_filter::containment_agent_type interleaved_bloom_filter::containment_agent() const
{
    return containment_agent_type{*this};
}

template <std::integral value_t>
inline interleaved_bloom_filter::counting_agent_type<value_t> interleaved_bloom_filter::counting_agent() const
{
    return counting_agent_type<value_t>{*this};
}

inline interleaved_bloom_filter::membership_agent_type interleaved_bloom_filter::membership_agent() const
{
    return membership_agent_type{*this};
}

} // namespace seqan::hibf
