// seqan::hibf::print: writes a range as "[a,b,c]\n" (src/misc/print.cpp), the
// format of the reference's snippet outputs.
#pragma once

#include <cstdint>
#include <iostream>
#include <ranges>
#include <type_traits>

namespace seqan::hibf
{

template <typename range_t>
void print(range_t const & range, std::ostream & os = std::cout)
{
    os << '[';
    bool first = true;
    // The element type decides the format, not the iterator's reference type:
    // a bit_vector yields bit proxies whose value type is bool, printed as 0/1
    // whatever the stream's boolalpha flag.
    using v_t = std::ranges::range_value_t<range_t>;
    for (auto && v : range)
    {
        if (!first)
            os << ',';
        first = false;
        if constexpr (std::is_same_v<v_t, bool> || sizeof(v_t) == 1)
            os << static_cast<int>(static_cast<v_t>(v));
        else
            os << static_cast<v_t>(v);
    }
    os << "]\n";
}

} // namespace seqan::hibf
