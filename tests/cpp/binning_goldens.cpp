// The layout DP against the expectations of the reference's
// test/unit/hibf/layout/hierarchical_binning_test.cpp:38-302 and
// simple_binning_test.cpp:17-72 (inputs and expected records copied as data),
// and the FPR corrections of fpr_correction_test.cpp:13-43.
// The hierarchical cases run with the tests' raw tmax (no rounding to 64) and
// union estimation disabled; compute_layout orders the merged-bin records by
// depth (stably), so the expected ones are ordered the same way before the
// comparison.  Prints "<case> ok" or "<case> FAILED" per case.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <utility>
#include <stdexcept>
#include <string>
#include <vector>

#include "layout.hpp"

using namespace hibf_amd::host;
using MaxBin = Layout::MaxBin;
using UserBin = Layout::UserBin;

static int failures = 0;

static bool same(std::vector<UserBin> const & a, std::vector<UserBin> const & b)
{
    if (a.size() != b.size())
        return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].previous_TB_indices != b[i].previous_TB_indices || a[i].storage_TB_id != b[i].storage_TB_id
            || a[i].number_of_technical_bins != b[i].number_of_technical_bins || a[i].idx != b[i].idx)
            return false;
    return true;
}

static bool same(std::vector<MaxBin> a, std::vector<MaxBin> const & got)
{
    std::stable_sort(a.begin(), a.end(), [](MaxBin const & x, MaxBin const & y)
                     { return x.previous_TB_indices.size() < y.previous_TB_indices.size(); });
    if (a.size() != got.size())
        return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (a[i].previous_TB_indices != got[i].previous_TB_indices || a[i].id != got[i].id)
            return false;
    return true;
}

static void hierarchical(char const * name, uint64_t tmax, double alpha, double empty, std::vector<uint64_t> counts,
                         uint64_t top, std::vector<MaxBin> max_bins, std::vector<UserBin> user_bins,
                         bool check_max_bins = true)
{
    Config c;
    c.number_of_user_bins = counts.size();
    c.tmax = tmax;
    c.alpha = alpha;
    c.empty_bin_fraction = empty;
    c.disable_estimate_union = true;
    c.disable_rearrangement = true;
    Layout const l = compute_layout(c, counts);
    bool const ok = l.top_level_max_bin_id == top && (!check_max_bins || same(max_bins, l.max_bins))
                 && same(user_bins, l.user_bins);
    std::printf("%s %s\n", name, ok ? "ok" : "FAILED");
    failures += !ok;
}

int main()
{
    hierarchical("small_example", 4, 1.2, 0.0, {500, 1000, 500, 500, 500, 500, 500, 500}, 3,
                 {{{2}, 22}, {{3}, 0}},
                 {{{}, 0, 1, 7}, {{}, 1, 1, 6}, {{2}, 0, 22, 3}, {{2}, 22, 21, 4}, {{2}, 43, 21, 5}, {{3}, 0, 42, 1},
                  {{3}, 42, 11, 0}, {{3}, 53, 11, 2}});
    hierarchical("small_example_with_empty_bins", 5, 1.2, 0.001, {500, 1000, 500, 500, 500, 500, 500, 500}, 3,
                 {{{2}, 0}, {{3}, 43}},
                 {{{}, 0, 1, 7}, {{}, 1, 1, 6}, {{2}, 0, 21, 3}, {{2}, 21, 21, 4}, {{2}, 42, 21, 5},
                  {{3}, 0, 43, 1}, {{3}, 43, 10, 0}, {{3}, 53, 10, 2}});
    hierarchical("another_example", 5, 1.2, 0.0, {50, 1000, 1000, 50, 5, 10, 10, 5}, 1,
                 {{{0, 0}, 45}, {{0}, 1}},
                 {{{0, 0}, 0, 45, 6}, {{0, 0}, 45, 13, 5}, {{0, 0}, 58, 3, 7}, {{0, 0}, 61, 3, 4}, {{0}, 1, 2, 0},
                  {{0}, 3, 2, 3}, {{}, 1, 2, 2}, {{}, 3, 2, 1}});
    hierarchical("high_level_max_bin_id_is_0", 4, 1.2, 0.0, {500, 500, 500, 500}, 0, {},
                 {{{}, 0, 1, 3}, {{}, 1, 1, 2}, {{}, 2, 1, 1}, {{}, 3, 1, 0}}, false);
    hierarchical("knuts_example", 5, 1.0, 0.0, {60, 600, 1000, 800, 800}, 1, {{{0}, 0}},
                 {{{0}, 0, 63, 1}, {{0}, 63, 1, 0}, {{}, 1, 1, 4}, {{}, 2, 1, 3}, {{}, 3, 2, 2}});
    hierarchical("four_level_hibf", 2, 1.2, 0.0, {11090, 5080, 3040, 1020, 510, 500}, 1,
                 {{{0, 0}, 33}, {{0, 1}, 0}, {{0}, 1}, {{1}, 0}},
                 {{{0, 0}, 0, 33, 4}, {{0, 0}, 33, 31, 5}, {{0, 1}, 0, 57, 2}, {{0, 1}, 57, 7, 3}, {{1}, 0, 53, 0},
                  {{1}, 53, 11, 1}});
    hierarchical("tb0_is_a_merged_bin", 2, 1.0, 0.0, {500, 500, 500, 500}, 0, {{{0}, 0}, {{1}, 0}},
                 {{{0}, 0, 32, 2}, {{0}, 32, 32, 3}, {{1}, 0, 32, 0}, {{1}, 32, 32, 1}});
    hierarchical("tb0_is_a_merged_bin_and_leads_to_recursive_call", 2, 1.0, 0.0,
                 {500, 500, 500, 500, 500, 500, 500, 500}, 0,
                 {{{0, 0}, 0}, {{0, 1}, 0}, {{0}, 0}, {{1, 0}, 0}, {{1, 1}, 0}, {{1}, 0}},
                 {{{0, 0}, 0, 32, 5}, {{0, 0}, 32, 32, 4}, {{0, 1}, 0, 32, 7}, {{0, 1}, 32, 32, 6}, {{1, 0}, 0, 32, 1},
                  {{1, 0}, 32, 32, 0}, {{1, 1}, 0, 32, 3}, {{1, 1}, 32, 32, 2}});

    {   // fpr_correction_test.cpp:13-43
        std::vector<double> const one = fpr_correction(0.05, 2, 8);
        bool ok = one.size() > 1u;
        for (uint64_t const v : {9123u, 123u, 12u, 87123u, 8123u, 4660u})
            ok = ok && static_cast<double>(v) == static_cast<double>(v) * one[1];
        std::vector<double> const c = fpr_correction(0.01, 5, 256);
        std::vector<std::pair<size_t, double>> const want{{1, 1.0},       {2, 1.192316},  {4, 1.412390},
                                                          {8, 1.664459},  {16, 1.953384}, {32, 2.284738},
                                                          {64, 2.664909}, {128, 3.101225}, {256, 3.602093}};
        ok = ok && c.size() == 257u;
        for (auto const & [i, v] : want)
            ok = ok && c.size() > i && std::abs(c[i] - v) <= 0.00001;
        for (size_t i = 1; ok && i < 256u; ++i)
            ok = c[i] <= c[i + 1];
        std::printf("fpr_correction %s\n", ok ? "ok" : "FAILED");
        failures += !ok;
    }
    std::vector<double> const ones(65, 1.0);
    {
        Layout l;
        uint64_t const max_bin = simple_binning_layout({100, 40, 20, 20}, {0, 1, 2, 3}, ones, 9, l);
        bool const ok = max_bin == 0u && same({{{}, 0, 1, 3}, {{}, 1, 1, 2}, {{}, 2, 2, 1}, {{}, 4, 5, 0}}, l.user_bins);
        std::printf("simple_small_example %s\n", ok ? "ok" : "FAILED");
        failures += !ok;
    }
    {
        Layout l;
        uint64_t const max_bin = simple_binning_layout({20, 20, 20, 20}, {0, 1, 2, 3}, ones, 4, l);
        bool const ok = max_bin == 0u && same({{{}, 0, 1, 3}, {{}, 1, 1, 2}, {{}, 2, 1, 1}, {{}, 3, 1, 0}}, l.user_bins);
        std::printf("simple_uniform_distribution %s\n", ok ? "ok" : "FAILED");
        failures += !ok;
    }
    {
        Layout l;
        bool threw = false;
        try
        {
            simple_binning_layout({100, 40, 20, 20}, {0, 1, 2, 3}, ones, 2, l);
        }
        catch (std::runtime_error const &)
        {
            threw = true;
        }
        std::printf("simple_too_few_technical_bins %s\n", threw ? "ok" : "FAILED");
        failures += !threw;
    }
    return failures ? 1 : 0;
}
