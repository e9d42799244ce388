// Test infrastructure: drives the reference's own robin_hood::unordered_flat_set
// (include/hibf/contrib/robin_hood.hpp of the seqan/hibf tree, the set its
// build chunks split user bins from: src/build/insert_into_ibf.cpp:28-47,
// src/hierarchical_interleaved_bloom_filter.cpp:57-171) through a script of
// set operations read from stdin, and prints each dumped set's iteration
// order.  Built by `make -C oracle ref` into oracle/_ref/ from the reference
// header where it lies; it writes the golden orders the oracle's and the
// library's restatements of that set are checked against.
//
// Script, one operation per line:
//   gen S seed set begin count    emplace synth(seed, set, begin + j), j < count
//   genmod S seed set count mod   emplace synth(seed, set, j % mod), j < count (duplicates)
//   vals S v1 v2 ...              emplace the given values
//   merge D S                     D.insert(S.begin(), S.end())  (update_parent_kmers)
//   clear S                       S.clear()
//   dump S full                   print "dump S size mask digest" and, with full = 1, the order
// Sets are named by small integers and created empty on first use.
#include <hibf/contrib/robin_hood.hpp>

#include <cinttypes>
#include <cstdio>
#include <iostream>
#include <map>
#include <sstream>
#include <string>

namespace
{

// The synthetic test-input stream (not reference code): value i of set `set`.
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

using set_t = robin_hood::unordered_flat_set<uint64_t>;

// Order-sensitive digest of an iteration order (FNV-1a over whole values).
uint64_t digest(set_t const & s)
{
    uint64_t h = 1469598103934665603ULL;
    for (uint64_t v : s)
    {
        h ^= v;
        h *= 1099511628211ULL;
    }
    return h;
}

} // namespace

int main()
{
    std::map<int, set_t> sets;
    std::string line;
    while (std::getline(std::cin, line))
    {
        std::istringstream in(line);
        std::string op;
        if (!(in >> op))
            continue;
        int a = 0;
        in >> a;
        set_t & s = sets[a];
        if (op == "gen")
        {
            uint64_t seed, set, begin, count;
            in >> seed >> set >> begin >> count;
            for (uint64_t j = 0; j < count; ++j)
                s.emplace(synth(seed, set, begin + j));
        }
        else if (op == "genmod")
        {
            uint64_t seed, set, count, mod;
            in >> seed >> set >> count >> mod;
            for (uint64_t j = 0; j < count; ++j)
                s.emplace(synth(seed, set, j % mod));
        }
        else if (op == "vals")
        {
            uint64_t v;
            while (in >> v)
                s.emplace(v);
        }
        else if (op == "merge")
        {
            int b = 0;
            in >> b;
            set_t const & src = sets[b];
            sets[a].insert(src.begin(), src.end());
        }
        else if (op == "clear")
        {
            s.clear();
        }
        else if (op == "dump")
        {
            int full = 0;
            in >> full;
            std::printf("dump %d %zu %zu %" PRIu64, a, s.size(), s.mask(), digest(s));
            if (full)
                for (uint64_t v : s)
                    std::printf(" %" PRIu64, v);
            std::printf("\n");
        }
        else
        {
            std::fprintf(stderr, "unknown operation: %s\n", op.c_str());
            return 2;
        }
    }
    return 0;
}
