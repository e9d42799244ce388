// Host-side HIBF construction support: config validation, FPR corrections,
// the hierarchical layout DP and the layout text format.  This is the part of
// the reference that stays on the host; it produces the tables the device
// query path consumes.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include <hibf/sketch/hyperloglog.hpp>
#include <hibf/sketch/toolbox.hpp>

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

// Estimated sizes of the unions of contiguous runs of one DP level's ordered
// user bins (toolbox::precompute_union_estimates_for and
// precompute_initial_union_estimates).
class UnionEstimates
{
public:
    virtual ~UnionEstimates() = default;
    // A DP level starts; positions are its user bins in DP order (sorted and
    // rearranged).  Returns false on failure (the layout then fails).
    virtual bool start(std::vector<uint64_t> const & positions) = 0;
    // prefix[j] for 1 <= j < positions.size(): estimate of the union of positions[0 .. j].
    virtual bool prefix(uint64_t * prefix) = 0;
    // For j = 1, 2, ... in increasing order within one level: a pointer to j
    // values, the k-th the estimate of the union of positions[k .. j]; valid
    // until the next call.  Null on failure.
    virtual uint64_t const * column(uint64_t j) = 0;
};

// The reference's sequential merges on the host, one column at a time.
class HostUnionEstimates final : public UnionEstimates
{
public:
    explicit HostUnionEstimates(std::vector<seqan::hibf::sketch::hyperloglog> const & sketches) : sketches_{sketches} {}
    bool start(std::vector<uint64_t> const & positions) override;
    bool prefix(uint64_t * prefix) override;
    uint64_t const * column(uint64_t j) override;

private:
    std::vector<seqan::hibf::sketch::hyperloglog> const & sketches_;
    std::vector<uint64_t> positions_;
    std::vector<uint64_t> column_;
};

// What the union-estimate and rearrangement modes need: the user bins' sketches
// (host copies), the DP's union estimates, and the merged-sketch estimates of
// the clustering (toolbox::cluster_bins).  Null members fall back to the host.
struct Sketching
{
    std::vector<seqan::hibf::sketch::hyperloglog> const * sketches{};
    UnionEstimates * unions{};
    seqan::hibf::sketch::toolbox::detail::union_source * pair_unions{};
    double union_seconds{};       // time in union estimates (filled by compute_layout)
    double rearrangement_seconds{}; // time in rearrange_bins
};

// layout/compute_layout.cpp:26-70 with hierarchical_binning / simple_binning.
// Without sketches (or with disable_estimate_union) merged-bin cardinalities
// are the sums of the user-bin cardinalities and user bins are not rearranged
// (the reference's disable_estimate_union layout); with them the DP uses union
// estimates and, unless disable_rearrangement, rearranges each level's user
// bins by similarity first, as the reference.
// simple_binning.cpp:19-116 on its own: user bins `positions` (in that order)
// over num_technical_bins bins with the given FPR corrections; appends the
// records to `layout` and returns the id of the fullest bin.
uint64_t simple_binning_layout(std::vector<uint64_t> const & kmer_counts,
                               std::vector<uint64_t> const & positions,
                               std::vector<double> const & fpr_correction,
                               uint64_t num_technical_bins,
                               Layout & layout);

// positions: the user bins to lay out (all of them when null).
Layout compute_layout(Config const & c,
                      std::vector<uint64_t> const & kmer_counts,
                      Sketching * sketching = nullptr,
                      std::vector<uint64_t> const * positions = nullptr);

} // namespace hibf_amd::host
