// seqan::hibf::counting_vector<T> (host side): per-bin counters that wrap
// modulo 2^(8*sizeof(T)) (misc/counting_vector.hpp:143-313 of the reference).
#pragma once

#include <algorithm>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

#include <hibf/misc/bit_vector.hpp>

namespace seqan::hibf
{

template <std::integral value_t>
class counting_vector : public std::vector<value_t>
{
    using base_t = std::vector<value_t>;

public:
    using base_t::base_t;

    counting_vector & operator+=(bit_vector const & bv)
    {
        for (size_t w = 0; w < bv.word_count(); ++w)
            for (uint64_t x = bv.data()[w]; x; x &= x - 1u)
            {
                size_t const b = w * 64u + static_cast<size_t>(__builtin_ctzll(x));
                if (b < this->size())
                    (*this)[b] = static_cast<value_t>((*this)[b] + 1);
            }
        return *this;
    }
    counting_vector & operator-=(bit_vector const & bv)
    {
        for (size_t w = 0; w < bv.word_count(); ++w)
            for (uint64_t x = bv.data()[w]; x; x &= x - 1u)
            {
                size_t const b = w * 64u + static_cast<size_t>(__builtin_ctzll(x));
                if (b < this->size())
                    (*this)[b] = static_cast<value_t>((*this)[b] - 1);
            }
        return *this;
    }
    counting_vector & operator+=(counting_vector const & rhs)
    {
        std::transform(this->begin(), this->end(), rhs.begin(), this->begin(), std::plus<value_t>());
        return *this;
    }
    counting_vector & operator-=(counting_vector const & rhs)
    {
        std::transform(this->begin(), this->end(), rhs.begin(), this->begin(), std::minus<value_t>());
        return *this;
    }
};

} // namespace seqan::hibf
