// seqan::hibf::print: writes a range as "[a,b,c]\n" (src/misc/print.cpp), the
// format of the reference's snippet outputs.
#pragma once

#include <cstdint>
#include <iostream>
#include <type_traits>

namespace seqan::hibf
{

template <typename range_t>
void print(range_t const & range, std::ostream & os = std::cout)
{
    os << '[';
    bool first = true;
    for (auto && v : range)
    {
        if (!first)
            os << ',';
        first = false;
        using v_t = std::remove_cvref_t<decltype(v)>;
        if constexpr (std::is_same_v<v_t, bool> || sizeof(v_t) == 1)
            os << static_cast<int>(v);
        else
            os << v;
    }
    os << "]\n";
}

} // namespace seqan::hibf
