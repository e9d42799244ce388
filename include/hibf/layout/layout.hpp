// seqan::hibf::layout::layout: the records of a layout file (the reference's
// include/hibf/layout/layout.hpp and src/layout/layout.cpp).  The text format
// is the reference's:
//
//   #TOP_LEVEL_IBF fullest_technical_bin_idx:<id>
//   #LOWER_LEVEL_IBF_<i;j;...> fullest_technical_bin_idx:<id>     (one per lower-level IBF)
//   #USER_BIN_IDX<TAB>TECHNICAL_BIN_INDICES<TAB>NUMBER_OF_TECHNICAL_BINS
//   <idx><TAB><merged TBs;...;storage TB><TAB><1;...;1;number of TBs>   (one per user bin)
//
// read_from stops before the next '#' line, so partitioned layouts (several
// layouts one after the other) read one at a time.
#pragma once

#include <charconv>
#include <compare>
#include <cstddef>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

#include <hibf/layout/prefixes.hpp>

namespace seqan::hibf::layout
{

struct layout
{
    struct max_bin
    {
        std::vector<size_t> previous_TB_indices{}; // merged bins leading to the IBF
        size_t id{};                               // its technical bin with the most k-mers

        friend auto operator<=>(max_bin const &, max_bin const &) = default;

        friend std::ostream & operator<<(std::ostream & stream, max_bin const & m)
        {
            stream << prefix::layout_header << prefix::layout_lower_level << '_';
            for (size_t i = 0; i < m.previous_TB_indices.size(); ++i)
                stream << (i ? ";" : "") << m.previous_TB_indices[i];
            return stream << ' ' << prefix::layout_fullest_technical_bin_idx << m.id;
        }
    };

    struct user_bin
    {
        std::vector<size_t> previous_TB_indices{}; // merged bins above the IBF that stores it
        size_t storage_TB_id{};                    // first technical bin of the user bin
        size_t number_of_technical_bins{};         // 1: single bin, > 1: split bin
        size_t idx{};                              // index of the user bin in the input

        friend auto operator<=>(user_bin const &, user_bin const &) = default;

        friend std::ostream & operator<<(std::ostream & stream, user_bin const & u)
        {
            stream << u.idx << '\t';
            for (size_t b : u.previous_TB_indices)
                stream << b << ';';
            stream << u.storage_TB_id << '\t';
            for (size_t i = 0; i < u.previous_TB_indices.size(); ++i)
                stream << "1;"; // one technical bin per merged level
            return stream << u.number_of_technical_bins;
        }
    };

    size_t top_level_max_bin_id{};
    std::vector<max_bin> max_bins{};
    std::vector<user_bin> user_bins{};

    bool operator==(layout const &) const = default;

    void clear()
    {
        top_level_max_bin_id = 0;
        max_bins.clear();
        user_bins.clear();
    }

    void write_to(std::ostream & stream) const
    {
        stream << prefix::layout_first_header_line << ' ' << prefix::layout_fullest_technical_bin_idx
               << top_level_max_bin_id << '\n';
        for (max_bin const & m : max_bins)
            stream << m << '\n';
        stream << prefix::layout_column_names << '\n';
        for (user_bin const & u : user_bins)
            stream << u << '\n';
    }

    void read_from(std::istream & stream)
    {
        std::string line;
        std::getline(stream, line); // #TOP_LEVEL_IBF fullest_technical_bin_idx:<id>
        top_level_max_bin_id = number_after(line, prefix::layout_fullest_technical_bin_idx);
        while (std::getline(stream, line) && line != prefix::layout_column_names)
        {
            // #LOWER_LEVEL_IBF_<i;j;...> fullest_technical_bin_idx:<id>
            size_t const first = prefix::layout_header.size() + prefix::layout_lower_level.size() + 1u;
            if (line.size() < first)
                malformed(line);
            size_t const space = line.find(' ', first);
            max_bins.push_back(max_bin{numbers(std::string_view{line}.substr(first, space - first)),
                                       number_after(line, prefix::layout_fullest_technical_bin_idx)});
        }
        while (stream.good() && stream.peek() != prefix::layout_header[0] && std::getline(stream, line))
        {
            // <idx>\t<b;...;storage>\t<1;...;count>
            size_t const t1 = line.find('\t');
            size_t const t2 = t1 == std::string::npos ? t1 : line.find('\t', t1 + 1);
            if (t2 == std::string::npos)
                malformed(line);
            std::vector<size_t> const idx = numbers(std::string_view{line}.substr(0, t1));
            std::vector<size_t> bins = numbers(std::string_view{line}.substr(t1 + 1, t2 - t1 - 1));
            std::vector<size_t> const counts = numbers(std::string_view{line}.substr(t2 + 1));
            if (idx.empty() || bins.empty() || counts.empty())
                malformed(line);
            user_bin u{};
            u.idx = idx.front();
            u.storage_TB_id = bins.back();
            bins.pop_back();
            u.previous_TB_indices = std::move(bins);
            u.number_of_technical_bins = counts.back();
            user_bins.push_back(std::move(u));
        }
    }

private:
    // The reference asserts the format (src/layout/layout.cpp:31-47) and reads out of bounds without the
    // assertions; here a line that does not have the format is an error.
    [[noreturn]] static void malformed(std::string const & line)
    {
        throw std::invalid_argument{"[HIBF LAYOUT ERROR] Malformed layout line: " + line.substr(0, 80)};
    }

    // The ';'-separated numbers of a field.
    static std::vector<size_t> numbers(std::string_view s)
    {
        std::vector<size_t> out;
        char const * p = s.data();
        char const * const end = s.data() + s.size();
        while (p < end)
        {
            size_t v{};
            auto const r = std::from_chars(p, end, v);
            if (r.ptr == p)
                break;
            out.push_back(v);
            p = r.ptr + 1; // skip ';'
        }
        return out;
    }

    static size_t number_after(std::string const & line, std::string_view key)
    {
        size_t const at = line.find(key);
        size_t v{};
        if (at != std::string::npos)
            std::from_chars(line.data() + at + key.size(), line.data() + line.size(), v);
        return v;
    }
};

} // namespace seqan::hibf::layout
