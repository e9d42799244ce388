// store/load of the seqan::hibf C++ drop-in classes: an IBF and an HIBF are
// stored (the HIBF appended after a one-byte header, as a wrapper format
// would), loaded back into fresh objects, and queried; every line prints
// "same" when the loaded object answers exactly as the original.
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <ranges>
#include <vector>

#include <hibf/config.hpp>
#include <hibf/hierarchical_interleaved_bloom_filter.hpp>
#include <hibf/interleaved_bloom_filter.hpp>

namespace sh = seqan::hibf;

int main(int argc, char ** argv)
{
    std::filesystem::path const dir = argc > 1 ? argv[1] : std::filesystem::temp_directory_path();
    std::filesystem::path const path = dir / "archive_roundtrip.bin";

    sh::interleaved_bloom_filter ibf{sh::bin_count{130u}, sh::bin_size{4096u}, sh::hash_function_count{3u}, true};
    for (uint64_t v = 0; v < 2000u; ++v)
        ibf.emplace(v * 7919u, sh::bin_index{v % 130u});
    ibf.occupancy[5] = 42u;

    std::vector<std::vector<uint64_t>> user_bins(40);
    for (size_t u = 0; u < user_bins.size(); ++u)
        for (uint64_t v = 0; v < 50u + 10u * u; ++v)
            user_bins[u].push_back(u * 1000003u + v * 17u);
    sh::config config{.input_fn = [&](size_t const u, sh::insert_iterator it)
                      {
                          for (uint64_t v : user_bins[u])
                              it = v;
                      },
                      .number_of_user_bins = user_bins.size(),
                      .tmax = 64u};
    sh::hierarchical_interleaved_bloom_filter hibf{config};

    {
        std::ofstream header{path, std::ios::binary};
        header.put('\x13');
    }
    ibf.store(path, true);
    hibf.store(path, true);

    sh::interleaved_bloom_filter ibf2;
    sh::hierarchical_interleaved_bloom_filter hibf2;
    uint64_t const after_ibf = ibf2.load(path, 1u);
    uint64_t const end = hibf2.load(path, after_ibf);
    std::cout << "offsets " << (end == std::filesystem::file_size(path) ? "same" : "differ") << '\n';
    std::cout << "ibf " << (ibf2 == ibf ? "same" : "differ") << '\n';
    std::cout << "occupancy "
              << (ibf2.occupancy == ibf.occupancy && ibf2.track_occupancy && ibf2.occupancy[5] == 42u ? "same"
                                                                                                     : "differ")
              << '\n';

    auto a = ibf.counting_agent<uint16_t>();
    auto b = ibf2.counting_agent<uint16_t>();
    bool counts_same = true;
    for (uint64_t start = 0; start < 2000u; start += 250u)
    {
        std::vector<uint64_t> q;
        for (uint64_t v = start; v < start + 250u; ++v)
            q.push_back(v * 7919u);
        auto const & x = a.bulk_count(q);
        std::vector<uint16_t> const xs(x.begin(), x.end());
        auto const & y = b.bulk_count(q);
        counts_same = counts_same && std::vector<uint16_t>(y.begin(), y.end()) == xs;
    }
    std::cout << "ibf counts " << (counts_same ? "same" : "differ") << '\n';

    bool tables_same = hibf2.number_of_user_bins == hibf.number_of_user_bins && hibf2.next_ibf_id == hibf.next_ibf_id
                    && hibf2.ibf_bin_to_user_bin_id == hibf.ibf_bin_to_user_bin_id;
    std::cout << "hibf tables " << (tables_same ? "same" : "differ") << '\n';
    auto m = hibf.membership_agent();
    auto n = hibf2.membership_agent();
    bool ids_same = true;
    for (size_t u = 0; u < user_bins.size(); ++u)
    {
        std::vector<uint64_t> const r1 = m.membership_for(user_bins[u], 30u);
        std::vector<uint64_t> const r2 = n.membership_for(user_bins[u], 30u);
        ids_same = ids_same && r1 == r2 && !r1.empty();
    }
    std::cout << "hibf ids " << (ids_same ? "same" : "differ") << '\n';
    std::filesystem::remove(path);
    return 0;
}
