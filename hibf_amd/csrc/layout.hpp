// Host-side HIBF construction support: config validation, FPR corrections,
// the hierarchical layout DP and the layout text format.  This is the part of
// the reference that stays on the host; it produces the tables the device
// query path consumes.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace hibf_amd::host
{

// Numeric fields of seqan::hibf::config (config.hpp:111-324); input_fn is
// replaced by explicit user-bin contents handed to the builder.
struct Config
{
    uint64_t number_of_user_bins{};
    uint64_t number_of_hash_functions{2};
    double maximum_fpr{0.05};
    double relaxed_fpr{0.3};
    uint64_t threads{1};
    uint64_t sketch_bits{12};
    uint64_t tmax{};
    double empty_bin_fraction{};
    double alpha{1.2};
    double max_rearrangement_ratio{0.5};
    bool disable_estimate_union{false};
    bool disable_rearrangement{false};
    bool track_occupancy{false};
};

// config.cpp:64-130 (minus the input_fn check).  Returns an empty string on
// success, else the reference's message.
std::string validate_and_set_defaults(Config & c);

// build/bin_size_in_bits.hpp:35-45
uint64_t bin_size_in_bits(double fpr, uint64_t hash_count, uint64_t elements);
// layout/compute_fpr_correction.cpp:18-35 (index = number of split bins)
std::vector<double> fpr_correction(double fpr, uint64_t hash_count, uint64_t t_max);
// layout/compute_relaxed_fpr_correction.cpp:17-31
double relaxed_fpr_correction(double fpr, double relaxed_fpr, uint64_t hash_count);
// misc/subtract_empty_bins.hpp:25-33
uint64_t subtract_empty_bins(uint64_t tmax, double fraction);
inline uint64_t next_multiple_of_64(uint64_t v) { return ((v + 63u) >> 6) << 6; }

// layout/layout.hpp:27-112
struct Layout
{
    struct MaxBin
    {
        std::vector<uint64_t> previous_TB_indices;
        uint64_t id{};
    };
    struct UserBin
    {
        std::vector<uint64_t> previous_TB_indices;
        uint64_t storage_TB_id{};
        uint64_t number_of_technical_bins{};
        uint64_t idx{};
    };
    uint64_t top_level_max_bin_id{};
    std::vector<MaxBin> max_bins;
    std::vector<UserBin> user_bins;

    std::string to_text() const;                              // layout.cpp write_to
    static bool from_text(std::string const & text, Layout & out); // layout.cpp read_from
};

// layout/compute_layout.cpp:26-70 with hierarchical_binning / simple_binning.
// Union cardinalities of merged bins are the sums of the user-bin
// cardinalities (the reference's disable_estimate_union mode), and user bins
// are not rearranged by similarity.
Layout compute_layout(Config const & c, std::vector<uint64_t> const & kmer_counts);

} // namespace hibf_amd::host
