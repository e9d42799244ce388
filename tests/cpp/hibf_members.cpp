// The public members of seqan::hibf::hierarchical_interleaved_bloom_filter
// beyond the agents (hierarchical_interleaved_bloom_filter.hpp:145-290):
// ibf_vector, prev_ibf_id, version and the build timers.  Each line prints
// "<check> ok" or "<check> FAILED".
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <ranges>
#include <sstream>
#include <string>
#include <vector>

#include <hibf/config.hpp>
#include <hibf/hierarchical_interleaved_bloom_filter.hpp>
#include <hibf/interleaved_bloom_filter.hpp>

namespace sh = seqan::hibf;
using hibf_t = sh::hierarchical_interleaved_bloom_filter;

static void report(char const * what, bool ok)
{
    std::cout << what << (ok ? " ok" : " FAILED") << '\n';
}

int main()
{
    std::vector<std::vector<uint64_t>> user_bins(300);
    for (size_t u = 0; u < user_bins.size(); ++u)
        for (uint64_t v = 0; v < 40u + (u * 37u) % 400u; ++v)
            user_bins[u].push_back(u * 1000003u + v * 17u);
    sh::config config{.input_fn = [&](size_t const u, sh::insert_iterator it)
                      {
                          for (uint64_t v : user_bins[u])
                              it = v;
                      },
                      .number_of_user_bins = user_bins.size(),
                      .tmax = 64u};
    hibf_t hibf{config};
    size_t const n = hibf.next_ibf_id.size();

    bool ok = hibf.ibf_vector.size() == n && hibf.prev_ibf_id.size() == n && n > 1;
    for (size_t i = 0; ok && i < n; ++i)
        ok = hibf.ibf_vector[i].bin_count() == hibf.next_ibf_id[i].size()
          && hibf.ibf_vector[i].hash_function_count() == 2u;
    report("ibf_vector_geometry", ok);

    ok = hibf.prev_ibf_id[0] == hibf_t::previous_ibf_id_pair{0u, 0u};
    size_t merged = 0;
    for (size_t i = 0; i < n; ++i)
        for (size_t b = 0; b < hibf.next_ibf_id[i].size(); ++b)
            if (hibf.ibf_bin_to_user_bin_id[i][b] == HIBF_GPU_BIN_MERGED)
            {
                ++merged;
                ok = ok && hibf.prev_ibf_id[hibf.next_ibf_id[i][b]] == hibf_t::previous_ibf_id_pair{i, b};
            }
    report("prev_ibf_id", ok && merged + 1 == n);

    // A user bin stored in IBF 0: its run of technical bins counts every one of its values (no false negatives).
    ok = false;
    for (size_t b = 0; b < hibf.ibf_bin_to_user_bin_id[0].size() && !ok; ++b)
    {
        uint64_t const u = hibf.ibf_bin_to_user_bin_id[0][b];
        if (u >= user_bins.size())
            continue;
        auto agent = hibf.ibf_vector[0].counting_agent<uint32_t>();
        auto const & counts = agent.bulk_count(user_bins[u]);
        uint64_t sum = 0;
        for (size_t c = 0; c < counts.size(); ++c)
            if (hibf.ibf_bin_to_user_bin_id[0][c] == u)
                sum += counts[c];
        ok = sum >= user_bins[u].size();
    }
    report("ibf_vector_counts", ok);

    hibf_t copy{hibf};
    ok = copy.ibf_vector.size() == n && copy.ibf_vector[0] == hibf.ibf_vector[0];
    {
        auto a = hibf.membership_agent();
        auto b = copy.membership_agent();
        std::vector<uint64_t> q(user_bins[7].begin(), user_bins[7].end());
        auto ra = a.membership_for(q, 10u);
        auto rb = b.membership_for(q, 10u);
        ok = ok && ra == rb && !ra.empty();
    }
    report("copy_shares_device_hibf", ok);

    report("bin_count_fixed", !hibf.ibf_vector[0].try_increase_bin_number_to(sh::bin_count{hibf.ibf_vector[0].bin_count() + 1u}));
    report("version", hibf_t::version == 1u);

    // interleaved_bloom_filter_test.cpp:209-239 (emplace_with_occupancy): 64 hashes into each of the
    // first 64 of 128 bins of an IBF that tracks occupancy; emplacing them again adds nothing.
    {
        sh::interleaved_bloom_filter ibf{sh::bin_count{128u}, sh::bin_size{512}, sh::hash_function_count{2u}, true};
        for (size_t bin_idx = 0; bin_idx < 64u; ++bin_idx)
            for (size_t hash = 0; hash < 64u; ++hash)
                ibf.emplace(hash, sh::bin_index{bin_idx});
        auto agent = ibf.containment_agent();
        bool occ_ok = ibf.occupancy.size() == 128u;
        for (size_t hash = 0; hash < 64u; ++hash)
        {
            auto const & res = agent.bulk_contains(hash);
            for (size_t b = 0; b < 128u; ++b)
                occ_ok = occ_ok && res[b] == (b < 64u);
        }
        for (size_t b = 0; b < 128u; ++b)
            occ_ok = occ_ok && ((b < 64u) == (ibf.occupancy[b] != 0u));
        std::vector<size_t> const first = ibf.occupancy;
        for (size_t bin_idx = 0; bin_idx < 64u; ++bin_idx)
            ibf.emplace_range(std::views::iota(size_t{0}, size_t{64}), sh::bin_index{bin_idx});
        report("emplace_with_occupancy", occ_ok && ibf.occupancy == first);
    }

    // hierarchical_interleaved_bloom_filter_test.cpp:59-106: config and layout read from one stream.
    {
        std::vector<std::vector<size_t>> hashes{{1u, 2u, 3u, 4u, 5u, 6u, 7u, 8u, 9u, 10u}, {1u, 2u, 3u, 4u, 5u}};
        std::stringstream stream{"@HIBF_CONFIG\n@{\n@    \"hibf_config\": {\n@        \"version\": 1,\n"
                                 "@        \"number_of_user_bins\": 2,\n@        \"number_of_hash_functions\": 2,\n"
                                 "@        \"maximum_fpr\": 0.05,\n@        \"relaxed_fpr\": 0.3,\n"
                                 "@        \"threads\": 1,\n@        \"sketch_bits\": 12,\n@        \"tmax\": 64,\n"
                                 "@        \"alpha\": 1.2,\n@        \"max_rearrangement_ratio\": 0.5,\n"
                                 "@        \"disable_estimate_union\": false,\n"
                                 "@        \"disable_rearrangement\": true\n@    }\n@}\n@HIBF_CONFIG_END\n"
                                 "#TOP_LEVEL_IBF fullest_technical_bin_idx:0\n"
                                 "#USER_BIN_IDX\tTECHNICAL_BIN_INDICES\tNUMBER_OF_TECHNICAL_BINS\n"
                                 "1\t0\t34\n0\t34\t30\n"};
        sh::config cfg{};
        cfg.read_from(stream);
        cfg.input_fn = [&](size_t const num, sh::insert_iterator it)
        {
            for (auto const h : hashes[num])
                it = h;
        };
        sh::layout::layout l{};
        l.read_from(stream);
        hibf_t h{cfg, l};
        auto agent = h.membership_agent();
        std::vector<size_t> query{1u, 2u, 3u, 4u, 5u};
        auto & result = agent.membership_for(query, 2u);
        agent.sort_results();
        report("build_from_layout", result == std::vector<uint64_t>{0u, 1u} && h.number_of_user_bins == 2u
                                        && l.user_bins.size() == 2u && l.user_bins[0].number_of_technical_bins == 34u);
    }
    report("timers", hibf.fill_ibf_timer.in_seconds() > 0.0 && hibf.index_allocation_timer.in_seconds() > 0.0
                         && hibf.layout_compute_sketches_timer.in_seconds() > 0.0
                         && hibf.layout_dp_algorithm_timer.in_seconds() > 0.0
                         && hibf.user_bin_io_timer.in_seconds() > 0.0
                         && hibf.layout_union_estimation_timer.in_seconds() > 0.0
                         && hibf.layout_rearrangement_timer.in_seconds() > 0.0);
    return 0;
}
