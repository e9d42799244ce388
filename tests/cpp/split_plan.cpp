// Host-only driver of the library's HIBF build plan (hibf_amd/csrc/build.cpp
// and layout.cpp, compiled with the host compiler; no device).
//
//   split_plan LAYOUT SIZES SEED THREADS TMAX [full]
//     plans the build of synthetic user bins (value i of user bin u =
//     synth(SEED, u, i), i < the u-th line of SIZES) on the layout in file
//     LAYOUT, replaying the reference's k-mer sets for the split user bins on
//     THREADS threads, and prints one line per split user bin: user bin,
//     first technical bin, values, a digest of its chunk numbers (the sum over
//     values i of (chunk + 1) * splitmix64(i), mod 2^64; with `full` also the
//     chunk numbers), then "seconds S".
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "build.hpp"

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

class Synth final : public hibf_amd::host::ValueSource
{
public:
    Synth(uint64_t seed, std::vector<uint64_t> const & sizes) : seed_{seed}, sizes_{sizes} {}
    uint64_t read(uint32_t u, uint64_t pos, uint64_t cap, uint64_t * vals, uint32_t * tags) const override
    {
        if (pos >= sizes_[u])
            return 0;
        uint64_t const k = std::min(cap, sizes_[u] - pos);
        for (uint64_t j = 0; j < k; ++j)
        {
            vals[j] = synth(seed_, u, pos + j);
            tags[j] = static_cast<uint32_t>(pos + j);
        }
        return k;
    }

private:
    uint64_t seed_;
    std::vector<uint64_t> const & sizes_;
};

std::string slurp(char const * path)
{
    std::ifstream in(path);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

} // namespace

int main(int argc, char ** argv)
{
    if (argc < 6)
    {
        std::fprintf(stderr, "usage: split_plan LAYOUT SIZES SEED THREADS TMAX [full]\n");
        return 2;
    }
    std::vector<uint64_t> sizes;
    {
        std::ifstream in(argv[2]);
        for (uint64_t s; in >> s;)
            sizes.push_back(s);
    }
    uint64_t const seed = std::stoull(argv[3]);
    bool const full = argc > 6 && std::string(argv[6]) == "full";
    hibf_amd::host::Config c{};
    c.number_of_user_bins = sizes.size();
    c.tmax = std::stoull(argv[5]);
    if (std::string const e = hibf_amd::host::validate_and_set_defaults(c); !e.empty())
    {
        std::fprintf(stderr, "config: %s\n", e.c_str());
        return 1;
    }
    hibf_amd::host::Layout layout;
    if (!hibf_amd::host::Layout::from_text(slurp(argv[1]), layout))
    {
        std::fprintf(stderr, "layout could not be parsed\n");
        return 1;
    }
    Synth src{seed, sizes};
    hibf_amd::host::BuildInput in;
    in.ub_size = sizes;
    in.values = &src;
    in.threads = static_cast<unsigned>(std::stoul(argv[4]));
    hibf_amd::host::Plan plan;
    auto const t0 = std::chrono::steady_clock::now();
    std::string const e = hibf_amd::host::build_plan(c, layout, in, plan);
    double const secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!e.empty())
    {
        std::fprintf(stderr, "plan: %s\n", e.c_str());
        return 1;
    }
    for (auto const & seg : plan.segments)
    {
        if (!(seg.flags & hibf_amd::host::kChunked))
            continue;
        // digest: sum over values i of (chunk + 1) * mix(i), mod 2^64 (mix: the splitmix64 finaliser)
        uint64_t h = 0;
        std::string ids;
        for (uint64_t i = 0; i < seg.count; ++i)
        {
            uint64_t const at = seg.begin + i;
            uint64_t const cid = plan.chunk_id_bytes == 1u
                                     ? plan.chunk_ids[at]
                                     : plan.chunk_ids[2u * at] | (uint64_t{plan.chunk_ids[2u * at + 1u]} << 8);
            uint64_t m = i + 0x9E3779B97F4A7C15ULL;
            m = (m ^ (m >> 30)) * 0xBF58476D1CE4E5B9ULL;
            m = (m ^ (m >> 27)) * 0x94D049BB133111EBULL;
            m ^= m >> 31;
            h += (cid + 1u) * m;
            if (full)
                ids += " " + std::to_string(cid);
        }
        std::printf("%u %u %" PRIu64 " %" PRIu64 "%s\n", seg.ub, seg.tb, seg.count, h, ids.c_str());
    }
    std::printf("seconds %.3f\n", secs);
    return 0;
}
