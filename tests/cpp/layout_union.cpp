// The layout DP's union-estimate mode against its sum mode: with union
// "estimates" that are the exact sums of the counts, the DP must place every
// user bin exactly as with disable_estimate_union (hierarchical_binning.cpp
// reads union_estimates[j_prime + 1] where the sum mode adds the counts).
// Only the merged bins' sizes in the maximum-bin trackers differ (they come
// from the sketches), so user bins and merged-bin paths are compared, not the
// maximum-bin ids.  Reads "n tmax seed" from argv; prints "ok" or the first
// difference.
//
// "sketched FILE n bits tmax" instead prints the layout text of the host path
// (union estimates and rearrangement by the reference's sequential merges)
// for the counts (n uint64) and registers (n x 2^bits bytes) in FILE.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "layout.hpp"

using namespace hibf_amd::host;

class SumUnions final : public UnionEstimates
{
public:
    explicit SumUnions(std::vector<uint64_t> const & counts) : counts_{counts} {}
    bool start(std::vector<uint64_t> const & positions) override
    {
        positions_ = positions;
        ++levels;
        return true;
    }
    bool prefix(uint64_t * prefix) override
    {
        uint64_t sum = counts_[positions_[0]];
        for (size_t j = 1; j < positions_.size(); ++j)
            prefix[j] = sum += counts_[positions_[j]];
        return true;
    }
    uint64_t const * column(uint64_t j) override
    {
        column_.assign(j, 0u);
        uint64_t sum = counts_[positions_[j]];
        for (uint64_t k = j; k-- > 0;)
            column_[k] = sum += counts_[positions_[k]];
        ++columns;
        return column_.data();
    }
    uint64_t levels{}, columns{};

private:
    std::vector<uint64_t> const & counts_;
    std::vector<uint64_t> positions_;
    std::vector<uint64_t> column_;
};

int sketched(char const * path, uint64_t n, uint32_t bits, uint64_t tmax)
{
    FILE * f = std::fopen(path, "rb");
    if (!f)
        return 2;
    std::vector<uint64_t> counts(n);
    std::vector<uint8_t> regs(n << bits);
    bool const ok = std::fread(counts.data(), 8, n, f) == n && std::fread(regs.data(), 1, regs.size(), f) == regs.size();
    std::fclose(f);
    if (!ok)
        return 2;
    std::vector<seqan::hibf::sketch::hyperloglog> sketches;
    for (uint64_t u = 0; u < n; ++u)
        sketches.push_back(seqan::hibf::sketch::hyperloglog::from_registers(static_cast<uint8_t>(bits),
                                                                            regs.data() + (u << bits)));
    Config c;
    c.number_of_user_bins = n;
    c.tmax = tmax;
    c.sketch_bits = bits;
    if (!validate_and_set_defaults(c).empty())
        return 2;
    Sketching sk;
    sk.sketches = &sketches;
    std::fputs(compute_layout(c, counts, &sk).to_text().c_str(), stdout);
    return 0;
}

int main(int argc, char ** argv)
{
    if (argc > 5 && std::string(argv[1]) == "sketched")
        return sketched(argv[2], std::strtoull(argv[3], nullptr, 10), static_cast<uint32_t>(std::atoi(argv[4])),
                        std::strtoull(argv[5], nullptr, 10));
    uint64_t const n = argc > 1 ? std::strtoull(argv[1], nullptr, 10) : 500;
    uint64_t const tmax = argc > 2 ? std::strtoull(argv[2], nullptr, 10) : 64;
    uint64_t const seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 1;
    std::mt19937_64 rng(seed);
    std::lognormal_distribution<double> size(std::log(2000.0), 1.5);
    std::vector<uint64_t> counts(n);
    for (auto & c : counts)
        c = 1 + static_cast<uint64_t>(size(rng));
    Config c;
    c.number_of_user_bins = n;
    c.tmax = tmax;
    c.disable_rearrangement = true;
    if (!validate_and_set_defaults(c).empty())
        return 2;
    Config off = c;
    off.disable_estimate_union = true;
    Layout const want = compute_layout(off, counts);

    std::vector<seqan::hibf::sketch::hyperloglog> sketches(n, seqan::hibf::sketch::hyperloglog{5});
    SumUnions unions{counts};
    Sketching sk;
    sk.sketches = &sketches;
    sk.unions = &unions;
    Layout const got = compute_layout(c, counts, &sk);

    if (got.user_bins.size() != want.user_bins.size() || got.max_bins.size() != want.max_bins.size())
    {
        std::printf("FAIL sizes %zu/%zu user bins, %zu/%zu max bins\n", got.user_bins.size(), want.user_bins.size(),
                    got.max_bins.size(), want.max_bins.size());
        return 1;
    }
    for (size_t i = 0; i < got.user_bins.size(); ++i)
    {
        auto const & a = got.user_bins[i];
        auto const & b = want.user_bins[i];
        if (a.previous_TB_indices != b.previous_TB_indices || a.storage_TB_id != b.storage_TB_id
            || a.number_of_technical_bins != b.number_of_technical_bins || a.idx != b.idx)
        {
            std::printf("FAIL user bin record %zu (ub %llu)\n", i, static_cast<unsigned long long>(a.idx));
            return 1;
        }
    }
    for (size_t i = 0; i < got.max_bins.size(); ++i)
        if (got.max_bins[i].previous_TB_indices != want.max_bins[i].previous_TB_indices)
        {
            std::printf("FAIL max bin record %zu\n", i);
            return 1;
        }
    std::printf("ok %zu user bins, %zu merged bins, %llu DP levels, %llu columns\n", got.user_bins.size(),
                got.max_bins.size(), static_cast<unsigned long long>(unions.levels),
                static_cast<unsigned long long>(unions.columns));
    return 0;
}
