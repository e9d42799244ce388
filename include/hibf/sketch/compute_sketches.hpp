// seqan::hibf::sketch::compute_sketches and estimate_kmer_counts
// (src/sketch/compute_sketches.cpp:27-43, include/hibf/sketch/estimate_kmer_counts.hpp):
// one hyperloglog of config.sketch_bits per user bin, here computed on the
// device for all user bins at once (hibf_gpu_hll_sketch) from the values that
// config.input_fn yields, and the truncated estimates of a list of sketches.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include <hibf/config.hpp>
#include <hibf/detail/device_buffer.hpp>
#include <hibf/misc/insert_iterator.hpp>
#include <hibf/sketch/hyperloglog.hpp>
#include <hibf_gpu.h>

namespace seqan::hibf::sketch
{

inline void compute_sketches(config const & config, std::vector<hyperloglog> & hll_sketches)
{
    size_t const n = config.number_of_user_bins;
    uint32_t const bits = config.sketch_bits;
    std::vector<uint64_t> values, offsets{0};
    std::vector<uint64_t> scratch;
    for (size_t i = 0; i < n; ++i)
    {
        scratch.clear();
        config.input_fn(i, insert_iterator{scratch});
        values.insert(values.end(), scratch.begin(), scratch.end());
        offsets.push_back(values.size());
    }
    detail::device_buffer d_values, d_offsets, d_regs;
    if (!values.empty())
        d_values.upload(values.data(), values.size() * sizeof(uint64_t));
    d_offsets.upload(offsets.data(), offsets.size() * sizeof(uint64_t));
    size_t const bytes = n << bits;
    d_regs.reserve(bytes ? bytes : 1u);
    detail::check(hibf_gpu_hll_sketch(d_values.as<uint64_t>(), d_offsets.as<uint64_t>(), n, bits,
                                      d_regs.as<uint8_t>(), nullptr));
    std::vector<uint8_t> regs(bytes);
    if (bytes)
        d_regs.download(regs.data(), bytes);
    hll_sketches.clear();
    hll_sketches.reserve(n);
    for (size_t i = 0; i < n; ++i)
        hll_sketches.push_back(hyperloglog::from_registers(static_cast<uint8_t>(bits), regs.data() + (i << bits)));
}

// estimate_kmer_counts.hpp: kmer_counts[i] = size_t(sketches[i].estimate()).
inline void estimate_kmer_counts(std::vector<hyperloglog> const & sketches, std::vector<size_t> & kmer_counts)
{
    kmer_counts.resize(sketches.size());
    for (size_t i = 0; i < sketches.size(); ++i)
        kmer_counts[i] = sketches[i].estimate();
}

} // namespace seqan::hibf::sketch
