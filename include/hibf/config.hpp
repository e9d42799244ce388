// seqan::hibf::config for the MI355X-native HIBF: same members, order and
// defaults as the reference (config.hpp:75-413), so designated-initialiser
// code compiles unchanged.  validate_and_set_defaults applies the reference's
// rules (src/config.cpp:64-130) through the C-ABI.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>

#include <hibf/misc/insert_iterator.hpp>
#include <hibf_gpu.h>

namespace seqan::hibf
{

struct config
{
    std::function<void(size_t const, insert_iterator &&)> input_fn{};
    size_t number_of_user_bins{};
    size_t number_of_hash_functions{2};
    double maximum_fpr{0.05};
    double relaxed_fpr{0.3};
    size_t threads{1u};
    uint8_t sketch_bits{12};
    size_t tmax{};
    double empty_bin_fraction{};
    bool track_occupancy{false};
    double alpha{1.2};
    double max_rearrangement_ratio{0.5};
    bool disable_estimate_union{false};
    bool disable_rearrangement{false};

    hibf_gpu_config to_c() const
    {
        hibf_gpu_config c{};
        hibf_gpu_config_default(&c);
        c.number_of_user_bins = number_of_user_bins;
        c.number_of_hash_functions = number_of_hash_functions;
        c.maximum_fpr = maximum_fpr;
        c.relaxed_fpr = relaxed_fpr;
        c.threads = threads;
        c.sketch_bits = sketch_bits;
        c.tmax = tmax;
        c.empty_bin_fraction = empty_bin_fraction;
        c.alpha = alpha;
        c.max_rearrangement_ratio = max_rearrangement_ratio;
        c.disable_estimate_union = disable_estimate_union;
        c.disable_rearrangement = disable_rearrangement;
        c.track_occupancy = track_occupancy;
        return c;
    }

    void validate_and_set_defaults()
    {
        if (!input_fn)
            throw std::invalid_argument{"[HIBF CONFIG ERROR] You did not set the required config::input_fn."};
        hibf_gpu_config c = to_c();
        if (hibf_gpu_config_validate(&c) != HIBF_GPU_OK)
            throw std::invalid_argument{hibf_gpu_last_error()};
        tmax = c.tmax;
        track_occupancy = c.track_occupancy != 0;
        disable_rearrangement = c.disable_rearrangement != 0;
    }

    bool operator==(config const & o) const
    {
        return number_of_user_bins == o.number_of_user_bins && number_of_hash_functions == o.number_of_hash_functions
            && maximum_fpr == o.maximum_fpr && relaxed_fpr == o.relaxed_fpr && threads == o.threads
            && sketch_bits == o.sketch_bits && tmax == o.tmax && empty_bin_fraction == o.empty_bin_fraction
            && alpha == o.alpha && max_rearrangement_ratio == o.max_rearrangement_ratio
            && disable_estimate_union == o.disable_estimate_union
            && disable_rearrangement == o.disable_rearrangement;
    }
};

} // namespace seqan::hibf
