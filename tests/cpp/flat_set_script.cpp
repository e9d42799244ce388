// Runs a golden-order script (the protocol of oracle/ref_robin_hood.cpp) on
// the library's FlatSetOrder (hibf_amd/csrc/flat_set_order.hpp), printing
// the same dump lines, so its iteration orders can be compared with the
// reference's robin_hood set and the oracle's restatement.
#include <cinttypes>
#include <cstdio>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "flat_set_order.hpp"

namespace
{

uint64_t synth(uint64_t seed, uint64_t set, uint64_t i)
{
    uint64_t k = ((set << 36) | i) + seed * 0x9E3779B97F4A7C15ULL;
    k ^= k >> 33;
    k *= 0xFF51AFD7ED558CCDULL;
    k ^= k >> 33;
    k *= 0xC4CEB9FE1A85EC53ULL;
    k ^= k >> 33;
    return k;
}

std::vector<uint64_t> order(hibf_amd::host::FlatSetOrder const & s)
{
    std::vector<uint64_t> out;
    out.reserve(s.size());
    s.for_each([&](uint64_t key, uint32_t) { out.push_back(key); });
    return out;
}

} // namespace

int main()
{
    std::map<int, hibf_amd::host::FlatSetOrder> sets;
    std::string line;
    while (std::getline(std::cin, line))
    {
        std::istringstream in(line);
        std::string op;
        int a = 0;
        if (!(in >> op >> a))
            continue;
        auto & s = sets[a];
        if (op == "gen")
        {
            uint64_t seed, set, begin, count;
            in >> seed >> set >> begin >> count;
            for (uint64_t j = 0; j < count; ++j)
                s.emplace(synth(seed, set, begin + j), static_cast<uint32_t>(j));
        }
        else if (op == "genmod")
        {
            uint64_t seed, set, count, mod;
            in >> seed >> set >> count >> mod;
            for (uint64_t j = 0; j < count; ++j)
                s.emplace(synth(seed, set, j % mod), static_cast<uint32_t>(j));
        }
        else if (op == "vals")
        {
            uint64_t v;
            while (in >> v)
                s.emplace(v, 0u);
        }
        else if (op == "merge")
        {
            int b = 0;
            in >> b;
            std::vector<uint64_t> const src = order(sets[b]);
            auto & dst = sets[a];
            for (uint64_t v : src)
                dst.emplace(v, 0u);
        }
        else if (op == "clear")
        {
            s.clear();
        }
        else if (op == "dump")
        {
            int full = 0;
            in >> full;
            std::vector<uint64_t> const o = order(s);
            uint64_t h = 1469598103934665603ULL;
            for (uint64_t v : o)
            {
                h ^= v;
                h *= 1099511628211ULL;
            }
            std::printf("dump %d %" PRIu64 " %" PRIu64 " %" PRIu64, a, s.size(), s.mask(), h);
            if (full)
                for (uint64_t v : o)
                    std::printf(" %" PRIu64, v);
            std::printf("\n");
        }
    }
    return 0;
}
