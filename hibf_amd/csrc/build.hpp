// Host-side HIBF build plan: the reference's hierarchical_build
// (src/hierarchical_interleaved_bloom_filter.cpp:40-179) restated as a plan
// of IBF geometries, routing tables and emplace segments that the device
// bulk emplace executes.
#pragma once

#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "layout.hpp"

namespace hibf_amd::host
{

// Sentinels of ibf_bin_to_user_bin_id (hierarchical_interleaved_bloom_filter.hpp:32-40).
constexpr uint64_t kBinMerged = 0xFFFFFFFFFFFFFFFFULL;
constexpr uint64_t kBinDeleted = 0xFFFFFFFFFFFFFFFEULL;

struct PlannedIbf
{
    uint64_t bins{};
    uint64_t bin_size{};
    uint64_t hash_funs{};
    std::vector<uint64_t> user_bin; // ibf_bin_to_user_bin_id[i]
    std::vector<uint64_t> next_ibf; // next_ibf_id[i]
};

// Values [begin, begin + count) of user bin `ub` (in its value order) go to
// technical bin `tb` of IBF `ibf`.
struct Segment
{
    uint32_t ibf;
    uint32_t tb;
    uint32_t ub;
    uint32_t pad;
    uint64_t begin;
    uint64_t count;
};

struct Plan
{
    std::vector<PlannedIbf> ibfs;
    std::vector<Segment> segments;
};

struct BuildInput
{
    std::vector<uint64_t> ub_size; // distinct values per user bin
    // Number of distinct values over a set of user bins (the size of the
    // merged bin's k-mer set); nullptr = user bins are pairwise disjoint.
    std::function<uint64_t(std::vector<uint32_t> const &)> union_size;
};

// Returns an empty string on success, else a description of the problem.
std::string build_plan(Config const & c, Layout const & layout, BuildInput const & in, Plan & out);

} // namespace hibf_amd::host
