// seqan::hibf::layout::compute_layout (include/hibf/layout/compute_layout.hpp,
// src/layout/compute_layout.cpp:26-70 in the reference): the hierarchical
// layout of the user bins from their k-mer counts and HyperLogLog sketches,
// with union estimates of merged bins and similarity rearrangement unless the
// config disables them.  The DP and the clustering run on the host inside the
// library; every union estimate they use is computed on the device
// (hibf_gpu_layout_compute_host_sketches).  Layouts are the reference's for
// the same inputs, with the clustering's one-thread order among equally
// distant pairs (see hibf/sketch/toolbox.hpp).
#pragma once

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include <hibf/config.hpp>
#include <hibf/detail/device_buffer.hpp>
#include <hibf/layout/layout.hpp>
#include <hibf/misc/timer.hpp>
#include <hibf/sketch/hyperloglog.hpp>
#include <hibf_gpu.h>

namespace seqan::hibf::layout
{

// compute_layout.cpp:26-55: only the user bins in `positions` are laid out;
// the union-estimation and rearrangement times are added to the timers.
inline layout compute_layout(config const & configuration,
                             std::vector<size_t> const & kmer_counts,
                             std::vector<sketch::hyperloglog> const & sketches,
                             std::vector<size_t> && positions,
                             concurrent_timer & union_estimation_timer,
                             concurrent_timer & rearrangement_timer)
{
    hibf_gpu_config c = configuration.to_c();
    detail::check(hibf_gpu_config_validate(&c));
    if (kmer_counts.size() != c.number_of_user_bins || sketches.size() != c.number_of_user_bins)
        throw std::invalid_argument{"[HIBF ERROR] kmer_counts and sketches must hold one entry per user bin."};
    uint64_t const m = uint64_t{1} << c.sketch_bits;
    std::vector<uint8_t> registers;
    registers.reserve(sketches.size() * m);
    for (sketch::hyperloglog const & s : sketches)
    {
        if (s.data_size() != m)
            throw std::invalid_argument{"[HIBF ERROR] every sketch must have config::sketch_bits bits."};
        registers.insert(registers.end(), s.registers().begin(), s.registers().end());
    }
    std::vector<uint64_t> const counts(kmer_counts.begin(), kmer_counts.end());
    std::vector<uint64_t> const pos(positions.begin(), positions.end());
    double times[2]{};
    uint64_t needed = 0;
    std::string text(std::max<size_t>(4096u, 64u * counts.size()), '\0');
    auto run = [&]()
    {
        return hibf_gpu_layout_compute_host_sketches(&c, counts.data(), registers.data(), pos.data(), pos.size(),
                                                     text.data(), text.size(), &needed, times);
    };
    int rc = run();
    if (rc == HIBF_GPU_ERR_CAPACITY)
    {
        text.assign(needed, '\0');
        rc = run();
    }
    detail::check(rc);
    union_estimation_timer.add_seconds(times[0]);
    rearrangement_timer.add_seconds(times[1]);
    layout result{};
    std::istringstream stream{std::string{text.c_str()}};
    result.read_from(stream);
    return result;
}

// compute_layout.cpp:57-70: every user bin.
inline layout compute_layout(config const & configuration,
                             std::vector<size_t> const & kmer_counts,
                             std::vector<sketch::hyperloglog> const & sketches)
{
    concurrent_timer union_estimation_timer;
    concurrent_timer rearrangement_timer;
    std::vector<size_t> positions(kmer_counts.size());
    for (size_t i = 0; i < positions.size(); ++i)
        positions[i] = i;
    return compute_layout(configuration, kmer_counts, sketches, std::move(positions), union_estimation_timer,
                          rearrangement_timer);
}

} // namespace seqan::hibf::layout
