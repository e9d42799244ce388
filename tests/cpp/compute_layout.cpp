// seqan::hibf::layout::compute_layout on the device: the reference's
// compute_layout_test.cpp:21-54 "dispatch" case (both overloads give the same
// layout), a layout of 300 user bins whose text is printed for comparison with
// the library's own sketched layout, and a layout of a subset of positions.
// Prints "<check> ok" or "<check> FAILED" per line, then the 300-bin layout.
#include <cstdio>
#include <iostream>
#include <random>
#include <sstream>
#include <vector>

#include <hibf/config.hpp>
#include <hibf/layout/compute_layout.hpp>
#include <hibf/sketch/compute_sketches.hpp>

using namespace seqan::hibf;

static void report(char const * what, bool ok)
{
    std::printf("%s %s\n", what, ok ? "ok" : "FAILED");
}

int main()
{
    {
        std::vector<std::vector<size_t>> hashes{{1u, 2u, 3u, 4u, 5u, 6u, 7u, 8u, 9u, 10u}, {1u, 2u, 3u, 4u, 5u}};
        config cfg{.input_fn = [&](size_t const num, insert_iterator it)
                   {
                       for (auto const h : hashes[num])
                           it = h;
                   },
                   .number_of_user_bins = 2};
        cfg.validate_and_set_defaults();
        std::vector<sketch::hyperloglog> sketches;
        std::vector<size_t> kmer_counts;
        sketch::compute_sketches(cfg, sketches);
        sketch::estimate_kmer_counts(sketches, kmer_counts);
        auto const layout1 = layout::compute_layout(cfg, kmer_counts, sketches);
        concurrent_timer union_timer{}, rearrangement_timer{};
        auto const layout2 = layout::compute_layout(cfg, kmer_counts, sketches, std::vector<size_t>{0u, 1u},
                                                    union_timer, rearrangement_timer);
        report("dispatch", layout1 == layout2 && layout1.user_bins.size() == 2u);
    }
    std::mt19937_64 rng{5};
    std::vector<std::vector<uint64_t>> ubs(300);
    for (auto & u : ubs)
    {
        u.resize(50 + rng() % 3000);
        for (auto & v : u)
            v = rng();
    }
    config cfg{.input_fn = [&](size_t const num, insert_iterator it)
               {
                   for (auto const h : ubs[num])
                       it = h;
               },
               .number_of_user_bins = ubs.size(),
               .tmax = 64};
    cfg.validate_and_set_defaults();
    std::vector<sketch::hyperloglog> sketches;
    std::vector<size_t> kmer_counts;
    sketch::compute_sketches(cfg, sketches);
    sketch::estimate_kmer_counts(sketches, kmer_counts);
    concurrent_timer union_timer{}, rearrangement_timer{};
    std::vector<size_t> all(ubs.size());
    for (size_t i = 0; i < all.size(); ++i)
        all[i] = i;
    auto const full = layout::compute_layout(cfg, kmer_counts, sketches, std::move(all), union_timer,
                                             rearrangement_timer);
    report("timers", union_timer.in_seconds() > 0.0 && rearrangement_timer.in_seconds() > 0.0);
    std::vector<size_t> subset;
    for (size_t i = 0; i < ubs.size(); i += 3)
        subset.push_back(i);
    size_t const n_subset = subset.size();
    auto const part = layout::compute_layout(cfg, kmer_counts, sketches, std::move(subset), union_timer,
                                             rearrangement_timer);
    bool subset_ok = part.user_bins.size() == n_subset;
    for (auto const & u : part.user_bins)
        subset_ok = subset_ok && u.idx % 3u == 0u;
    report("positions_subset", subset_ok);
    std::ostringstream text;
    full.write_to(text);
    std::cout << text.str();
    return 0;
}
