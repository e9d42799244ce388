// seqan::hibf::bit_vector (host side): the result type of
// containment_agent_type::bulk_contains.  LSB-first bits in 64-bit words as
// the reference's bit_vector (misc/bit_vector.hpp:64-67, 1045-1054).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

namespace seqan::hibf
{

class bit_vector
{
public:
    bit_vector() = default;
    explicit bit_vector(size_t const bits, bool const value = false) :
        words_((bits + 63u) / 64u, value ? ~0ULL : 0ULL),
        size_{bits}
    {}

    size_t size() const noexcept { return size_; }
    bool empty() const noexcept { return size_ == 0; }
    bool operator[](size_t const i) const noexcept { return (words_[i >> 6] >> (i & 63u)) & 1u; }
    uint64_t * data() noexcept { return words_.data(); }
    uint64_t const * data() const noexcept { return words_.data(); }
    size_t word_count() const noexcept { return words_.size(); }

    bool all() const noexcept
    {
        for (size_t i = 0; i < size_; ++i)
            if (!(*this)[i])
                return false;
        return true;
    }
    bool any() const noexcept
    {
        for (size_t i = 0; i < size_; ++i)
            if ((*this)[i])
                return true;
        return false;
    }
    bool none() const noexcept { return !any(); }

    struct const_iterator
    {
        bit_vector const * v;
        size_t i;
        bool operator*() const noexcept { return (*v)[i]; }
        const_iterator & operator++() noexcept { ++i; return *this; }
        bool operator!=(const_iterator const & o) const noexcept { return i != o.i; }
        bool operator==(const_iterator const & o) const noexcept { return i == o.i; }
        bool operator<(const_iterator const & o) const noexcept { return i < o.i; }
    };
    const_iterator begin() const noexcept { return {this, 0}; }
    const_iterator end() const noexcept { return {this, size_}; }

private:
    std::vector<uint64_t> words_;
    size_t size_{};
};

} // namespace seqan::hibf
