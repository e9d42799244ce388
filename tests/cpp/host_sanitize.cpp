// Host code under the sanitizers: layout text and config text parsing, the
// build plan (with the replay of the reference's k-mer sets) and archive
// parsing, fed the files named on the command line.
//   host_sanitize layout FILE...   parse each layout and plan 4 user bins of 100 values (tmax 64)
//   host_sanitize files FILE...    read each file as config + layout (seqan::hibf::config::read_from,
//                                  layout::layout::read_from)
//   host_sanitize ibf FILE...      hibf_gpu_ibf_load of each file
//   host_sanitize hibf FILE...     hibf_gpu_hibf_load of each file
// Prints one status line per file; the sanitizers abort on any memory or
// undefined-behaviour error.
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include <hibf/config.hpp>
#include <hibf/layout/layout.hpp>

#include "build.hpp"
#include "hibf_gpu.h"

namespace
{

std::string slurp(char const * path)
{
    std::ifstream in(path, std::ios::binary);
    std::stringstream ss;
    ss << in.rdbuf();
    return ss.str();
}

class Synth final : public hibf_amd::host::ValueSource
{
public:
    uint64_t read(uint32_t u, uint64_t pos, uint64_t cap, uint64_t * vals, uint32_t * tags) const override
    {
        uint64_t const n = 100;
        if (pos >= n)
            return 0;
        uint64_t const k = std::min(cap, n - pos);
        for (uint64_t j = 0; j < k; ++j)
        {
            vals[j] = (uint64_t{u} << 32 | (pos + j)) * 0x9E3779B97F4A7C15ULL;
            tags[j] = static_cast<uint32_t>(pos + j);
        }
        return k;
    }
};

} // namespace

int main(int argc, char ** argv)
{
    if (argc < 2)
        return 2;
    std::string const mode = argv[1];
    for (int i = 2; i < argc; ++i)
    {
        std::string status;
        if (mode == "layout")
        {
            hibf_amd::host::Config c{};
            c.number_of_user_bins = 4;
            c.tmax = 64;
            (void)hibf_amd::host::validate_and_set_defaults(c);
            hibf_amd::host::Layout layout;
            if (!hibf_amd::host::Layout::from_text(slurp(argv[i]), layout))
                status = "unparsed";
            else
            {
                Synth src;
                hibf_amd::host::BuildInput in;
                in.ub_size.assign(4, 100);
                in.values = &src;
                in.threads = 2;
                hibf_amd::host::Plan plan;
                std::string const e = hibf_amd::host::build_plan(c, layout, in, plan);
                status = e.empty() ? "ok" : e;
            }
        }
        else if (mode == "files")
        {
            try
            {
                std::istringstream in(slurp(argv[i]));
                seqan::hibf::config c{};
                c.read_from(in);
                seqan::hibf::layout::layout l{};
                l.read_from(in);
                status = "ok";
            }
            catch (std::exception const & e)
            {
                status = std::string("exception: ") + e.what();
            }
        }
        else if (mode == "ibf" || mode == "hibf")
        {
            uint64_t end = 0;
            int rc;
            if (mode == "ibf")
            {
                hibf_gpu_ibf h = nullptr;
                rc = hibf_gpu_ibf_load(argv[i], 0, &h, &end);
            }
            else
            {
                hibf_gpu_hibf h = nullptr;
                rc = hibf_gpu_hibf_load(argv[i], 0, &h, &end);
            }
            status = std::to_string(rc) + " " + hibf_gpu_last_error();
        }
        std::printf("%s: %s\n", argv[i], status.c_str());
    }
    return 0;
}
