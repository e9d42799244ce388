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

// Values [begin, begin + count) of user bin `ub` (in its device order) go to
// technical bin `tb` of IBF `ibf`.  A chunked segment (kChunked in flags,
// the number of technical bins in flags >> 8) holds a whole split user bin:
// its value i goes to technical bin tb + chunk, the chunk being entry
// begin + i of Plan::chunk_ids.
struct Segment
{
    uint32_t ibf;
    uint32_t tb;
    uint32_t ub;
    uint32_t flags;
    uint64_t begin;
    uint64_t count;
};
constexpr uint32_t kChunked = 1u;

struct Plan
{
    std::vector<PlannedIbf> ibfs;
    std::vector<Segment> segments;
    // Chunk numbers of the split user bins' values (chunked segments), of
    // chunk_id_bytes (1 or 2) each.
    std::vector<uint8_t> chunk_ids;
    uint32_t chunk_id_bytes{1};
};

// The values of the user bins in the order the caller's input_fn would give
// them to the reference's build (duplicates included).
class ValueSource
{
public:
    virtual ~ValueSource() = default;
    // Writes up to cap values of user bin u from input position pos on into
    // vals, and each value's index in the device copy of the user bin into
    // tags; returns how many were written (0: past the end).
    virtual uint64_t read(uint32_t u, uint64_t pos, uint64_t cap, uint64_t * vals, uint32_t * tags) const = 0;
};

struct BuildInput
{
    std::vector<uint64_t> ub_size; // distinct values per user bin
    // Number of distinct values over a set of user bins (the size of the
    // merged bin's k-mer set); nullptr = user bins are pairwise disjoint.
    std::function<uint64_t(std::vector<uint32_t> const &)> union_size;
    // With a source, split user bins are chunked in the iteration order of the
    // reference's k-mer set (src/build/insert_into_ibf.cpp:28-47), replayed on
    // `threads` host threads; without one, in device order (ranges of values).
    ValueSource const * values{nullptr};
    unsigned threads{1};
};

// Returns an empty string on success, else a description of the problem.
std::string build_plan(Config const & c, Layout const & layout, BuildInput const & in, Plan & out);

} // namespace hibf_amd::host
