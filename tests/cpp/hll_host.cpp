// seqan::hibf::sketch::hyperloglog (host class) as a command-line filter for
// the tests: reads "bits n v1 ... vn" per line from stdin and prints the
// sketch's registers (hex) and estimate (%.17g) per line.  With the argument
// "device" it also builds the same sketches with compute_sketches (device)
// and prints "device same" when every register row is identical.
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include <hibf/sketch/hyperloglog.hpp>
#ifdef WITH_DEVICE
#include <hibf/sketch/compute_sketches.hpp>
#endif

int main(int argc, char ** argv)
{
    std::vector<std::vector<uint64_t>> sets;
    std::vector<unsigned> widths;
    std::string line;
    while (std::getline(std::cin, line))
    {
        std::istringstream in{line};
        unsigned bits{};
        size_t n{};
        in >> bits >> n;
        std::vector<uint64_t> v(n);
        for (auto & x : v)
            in >> x;
        seqan::hibf::sketch::hyperloglog h{static_cast<uint8_t>(bits)};
        for (uint64_t x : v)
            h.add(x);
        for (uint8_t r : h.registers())
            std::printf("%02x", r);
        std::printf(" %.17g\n", h.estimate());
        sets.push_back(std::move(v));
        widths.push_back(bits);
    }
#ifdef WITH_DEVICE
    if (argc > 1 && std::string{argv[1]} == "device" && !sets.empty())
    {
        seqan::hibf::config cfg{.input_fn = [&](size_t const u, seqan::hibf::insert_iterator it)
                                {
                                    for (uint64_t x : sets[u])
                                        it = x;
                                },
                                .number_of_user_bins = sets.size(),
                                .sketch_bits = static_cast<uint8_t>(widths[0])};
        std::vector<seqan::hibf::sketch::hyperloglog> sk;
        seqan::hibf::sketch::compute_sketches(cfg, sk);
        bool same = sk.size() == sets.size();
        for (size_t i = 0; same && i < sets.size(); ++i)
        {
            seqan::hibf::sketch::hyperloglog h{static_cast<uint8_t>(widths[0])};
            for (uint64_t x : sets[i])
                h.add(x);
            same = sk[i] == h && sk[i].estimate() == h.estimate();
        }
        std::vector<size_t> counts;
        seqan::hibf::sketch::estimate_kmer_counts(sk, counts);
        std::printf("device %s %zu\n", same ? "same" : "differ", counts.size());
    }
#else
    (void)argc;
    (void)argv;
#endif
    return 0;
}
