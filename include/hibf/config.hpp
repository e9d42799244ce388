// seqan::hibf::config for the MI355X-native HIBF: same members, order and
// defaults as the reference (config.hpp:75-413), so designated-initialiser
// code compiles unchanged.  validate_and_set_defaults applies the reference's
// rules (src/config.cpp:64-130) through the C-ABI.
#pragma once

#include <algorithm>
#include <charconv>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <istream>
#include <map>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include <hibf/misc/insert_iterator.hpp>
#include <hibf_gpu.h>

namespace seqan::hibf
{

namespace detail
{
// A double as a JSON number the way the reference's config files spell it
// (cereal's JSON archive): the shortest digits that read back to the same
// value; plain decimals for decimal exponents -6 < e <= 21, with ".0" after
// an integer; otherwise d[.ddd]e<exp> with no '+' and no leading zeros.
inline std::string json_number(double const x)
{
    char buf[64];
    auto const r = std::to_chars(buf, buf + sizeof(buf), x, std::chars_format::scientific);
    std::string sci(buf, r.ptr);
    std::string sign;
    if (!sci.empty() && sci[0] == '-')
    {
        sign = "-";
        sci.erase(0, 1);
    }
    size_t const e_pos = sci.find('e');
    std::string digits = sci.substr(0, e_pos);
    digits.erase(std::remove(digits.begin(), digits.end(), '.'), digits.end());
    int const exp10 = std::stoi(sci.substr(e_pos + 1));
    int const n = static_cast<int>(digits.size());
    int const kk = exp10 + 1;  // value = 0.digits x 10^kk
    int const k = kk - n;      // value = digits x 10^k
    std::string out;
    if (k >= 0 && kk <= 21)
        out = digits + std::string(static_cast<size_t>(k), '0') + ".0";
    else if (kk > 0 && kk <= 21)
        out = digits.substr(0, static_cast<size_t>(kk)) + "." + digits.substr(static_cast<size_t>(kk));
    else if (kk > -6 && kk <= 0)
        out = "0." + std::string(static_cast<size_t>(-kk), '0') + digits;
    else
        out = digits.substr(0, 1) + (n > 1 ? "." + digits.substr(1) : std::string{}) + "e" + std::to_string(kk - 1);
    return sign + out;
}
} // namespace detail

struct config
{
    std::function<void(size_t const, insert_iterator &&)> input_fn{};
    size_t number_of_user_bins{};
    size_t number_of_hash_functions{2};
    double maximum_fpr{0.05};
    double relaxed_fpr{0.3};
    size_t threads{1u};
    uint8_t sketch_bits{12};
    size_t tmax{};
    double empty_bin_fraction{};
    bool track_occupancy{false};
    double alpha{1.2};
    double max_rearrangement_ratio{0.5};
    bool disable_estimate_union{false};
    bool disable_rearrangement{false};

    hibf_gpu_config to_c() const
    {
        hibf_gpu_config c{};
        hibf_gpu_config_default(&c);
        c.number_of_user_bins = number_of_user_bins;
        c.number_of_hash_functions = number_of_hash_functions;
        c.maximum_fpr = maximum_fpr;
        c.relaxed_fpr = relaxed_fpr;
        c.threads = threads;
        c.sketch_bits = sketch_bits;
        c.tmax = tmax;
        c.empty_bin_fraction = empty_bin_fraction;
        c.alpha = alpha;
        c.max_rearrangement_ratio = max_rearrangement_ratio;
        c.disable_estimate_union = disable_estimate_union;
        c.disable_rearrangement = disable_rearrangement;
        c.track_occupancy = track_occupancy;
        return c;
    }

    void validate_and_set_defaults()
    {
        if (!input_fn)
            throw std::invalid_argument{"[HIBF CONFIG ERROR] You did not set the required config::input_fn."};
        hibf_gpu_config c = to_c();
        if (hibf_gpu_config_validate(&c) != HIBF_GPU_OK)
            throw std::invalid_argument{hibf_gpu_last_error()};
        tmax = c.tmax;
        track_occupancy = c.track_occupancy != 0;
        disable_rearrangement = c.disable_rearrangement != 0;
    }

    // config.hpp:380: version of the serialized form written by write_to.
    static constexpr uint32_t version{3u};

    // src/config.cpp:22-63: the "@HIBF_CONFIG" block of a layout file, one JSON document
    // ({"hibf_config": {...}} with the fields in the order of config.hpp:383-410) behind '@'.
    void write_to(std::ostream & stream) const
    {
        auto const b = [](bool v) { return std::string{v ? "true" : "false"}; };
        std::pair<char const *, std::string> const fields[] = {
            {"version", std::to_string(version)},
            {"number_of_user_bins", std::to_string(number_of_user_bins)},
            {"number_of_hash_functions", std::to_string(number_of_hash_functions)},
            {"maximum_fpr", detail::json_number(maximum_fpr)},
            {"relaxed_fpr", detail::json_number(relaxed_fpr)},
            {"threads", std::to_string(threads)},
            {"sketch_bits", std::to_string(static_cast<unsigned>(sketch_bits))},
            {"tmax", std::to_string(tmax)},
            {"empty_bin_fraction", detail::json_number(empty_bin_fraction)},
            {"track_occupancy", b(track_occupancy)},
            {"alpha", detail::json_number(alpha)},
            {"max_rearrangement_ratio", detail::json_number(max_rearrangement_ratio)},
            {"disable_estimate_union", b(disable_estimate_union)},
            {"disable_rearrangement", b(disable_rearrangement)}};
        stream << "@HIBF_CONFIG\n@{\n@    \"hibf_config\": {\n";
        size_t const n = std::size(fields);
        for (size_t i = 0; i < n; ++i)
            stream << "@        \"" << fields[i].first << "\": " << fields[i].second << (i + 1 < n ? ",\n" : "\n");
        stream << "@    }\n@}\n@HIBF_CONFIG_END\n";
    }

    // Reads the block written by write_to (versions 1-3) and leaves the stream after its last line.
    // Version 2 has no track_occupancy; it follows from empty_bin_fraction != 0 (config.hpp:396-403).
    void read_from(std::istream & stream)
    {
        std::string line;
        while (std::getline(stream, line) && line != "@HIBF_CONFIG")
            ;
        std::string json;
        while (std::getline(stream, line) && line != "@HIBF_CONFIG_END")
            json += line.substr(line.empty() ? 0 : 1) + '\n';
        std::map<std::string, std::string> kv;
        for (size_t p = json.find('"'); p != std::string::npos; p = json.find('"', p))
        {
            size_t const q = json.find('"', p + 1);
            if (q == std::string::npos)
                break;
            std::string const key = json.substr(p + 1, q - p - 1);
            size_t v = json.find_first_not_of(" \t\n:", q + 1);
            if (v == std::string::npos)
                break;
            if (json[v] == '{')
            {
                p = v + 1;
                continue;
            }
            size_t const end = json.find_first_of(",}\n", v);
            kv[key] = json.substr(v, end - v);
            p = end;
        }
        auto const u = [&](char const * k, auto & dst)
        {
            if (auto it = kv.find(k); it != kv.end())
                dst = static_cast<std::remove_reference_t<decltype(dst)>>(std::stoull(it->second));
        };
        auto const d = [&](char const * k, double & dst)
        {
            if (auto it = kv.find(k); it != kv.end())
                dst = std::stod(it->second);
        };
        auto const f = [&](char const * k, bool & dst)
        {
            if (auto it = kv.find(k); it != kv.end())
                dst = it->second.rfind("true", 0) == 0;
        };
        uint32_t parsed_version{version};
        u("version", parsed_version);
        u("number_of_user_bins", number_of_user_bins);
        u("number_of_hash_functions", number_of_hash_functions);
        d("maximum_fpr", maximum_fpr);
        d("relaxed_fpr", relaxed_fpr);
        u("threads", threads);
        u("sketch_bits", sketch_bits);
        u("tmax", tmax);
        if (parsed_version >= 2u)
        {
            d("empty_bin_fraction", empty_bin_fraction);
            if (parsed_version >= 3u)
                f("track_occupancy", track_occupancy);
            else
                track_occupancy = empty_bin_fraction != 0.0;
        }
        d("alpha", alpha);
        d("max_rearrangement_ratio", max_rearrangement_ratio);
        f("disable_estimate_union", disable_estimate_union);
        f("disable_rearrangement", disable_rearrangement);
    }

    bool operator==(config const & o) const
    {
        return number_of_user_bins == o.number_of_user_bins && number_of_hash_functions == o.number_of_hash_functions
            && maximum_fpr == o.maximum_fpr && relaxed_fpr == o.relaxed_fpr && threads == o.threads
            && sketch_bits == o.sketch_bits && tmax == o.tmax && empty_bin_fraction == o.empty_bin_fraction
            && alpha == o.alpha && max_rearrangement_ratio == o.max_rearrangement_ratio
            && disable_estimate_union == o.disable_estimate_union
            && disable_rearrangement == o.disable_rearrangement;
    }
};

} // namespace seqan::hibf
