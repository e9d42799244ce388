// seqan::hibf::config::write_to / read_from as a command-line filter for the
// tests: "write key=value ..." prints the "@HIBF_CONFIG" block of a config with
// those fields set; "read" parses such a block from stdin (after any other
// lines) and prints "key value" per field, then the rest of the input.
#include <cstdio>
#include <iostream>
#include <string>

#include <hibf/config.hpp>

int main(int argc, char ** argv)
{
    seqan::hibf::config c;
    std::string const mode = argc > 1 ? argv[1] : "";
    if (mode == "write")
    {
        for (int i = 2; i < argc; ++i)
        {
            std::string const kv = argv[i];
            std::string const k = kv.substr(0, kv.find('='));
            std::string const v = kv.substr(kv.find('=') + 1);
            if (k == "number_of_user_bins") c.number_of_user_bins = std::stoull(v);
            else if (k == "number_of_hash_functions") c.number_of_hash_functions = std::stoull(v);
            else if (k == "maximum_fpr") c.maximum_fpr = std::stod(v);
            else if (k == "relaxed_fpr") c.relaxed_fpr = std::stod(v);
            else if (k == "threads") c.threads = std::stoull(v);
            else if (k == "sketch_bits") c.sketch_bits = static_cast<uint8_t>(std::stoul(v));
            else if (k == "tmax") c.tmax = std::stoull(v);
            else if (k == "empty_bin_fraction") c.empty_bin_fraction = std::stod(v);
            else if (k == "track_occupancy") c.track_occupancy = v == "true";
            else if (k == "alpha") c.alpha = std::stod(v);
            else if (k == "max_rearrangement_ratio") c.max_rearrangement_ratio = std::stod(v);
            else if (k == "disable_estimate_union") c.disable_estimate_union = v == "true";
            else if (k == "disable_rearrangement") c.disable_rearrangement = v == "true";
            else return 2;
        }
        c.write_to(std::cout);
        return 0;
    }
    if (mode == "read")
    {
        c.read_from(std::cin);
        std::printf("number_of_user_bins %zu\nnumber_of_hash_functions %zu\nmaximum_fpr %.17g\nrelaxed_fpr %.17g\n"
                    "threads %zu\nsketch_bits %u\ntmax %zu\nempty_bin_fraction %.17g\ntrack_occupancy %d\n"
                    "alpha %.17g\nmax_rearrangement_ratio %.17g\ndisable_estimate_union %d\n"
                    "disable_rearrangement %d\n",
                    c.number_of_user_bins, c.number_of_hash_functions, c.maximum_fpr, c.relaxed_fpr, c.threads,
                    static_cast<unsigned>(c.sketch_bits), c.tmax, c.empty_bin_fraction, c.track_occupancy ? 1 : 0,
                    c.alpha, c.max_rearrangement_ratio, c.disable_estimate_union ? 1 : 0,
                    c.disable_rearrangement ? 1 : 0);
        std::string line;
        while (std::getline(std::cin, line))
            std::printf("rest %s\n", line.c_str());
        return 0;
    }
    return 2;
}
