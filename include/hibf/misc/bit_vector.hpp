// seqan::hibf::bit_vector (host side): the result type of
// containment_agent_type::bulk_contains, with the interface of the
// reference's include/hibf/misc/bit_vector.hpp:
//   - bits LSB-first in 64-bit words, word i holding bits 64i..64i+63
//     (:1045-1054), in a 64-byte aligned vector (:64-67);
//   - `reference` / `const_reference`: a proxy for one bit that converts to
//     bool; the writable one is assigned from bool or from another proxy and
//     flips (:90-178);
//   - random-access iterators over the bits; `iterator` is an output
//     iterator of bool, `const_iterator` is not (:184-360);
//   - construction from a count (and a bit) or an initializer list, assign,
//     back, size/empty/capacity/reserve, push_back, resize, clear, swap
//     (:400-530, 576-715, 821-830);
//   - word-wise all/any/none, &=, |=, ^=, ~, &, |, ^, and_not and flip, and
//     flip(position) that throws std::out_of_range past the end (:592-616,
//     717-819).
//
// The reference's word-wise behaviour is kept, since callers can observe it:
// all(), any() and none() test whole words, the bits past size() in the last
// word included; shrinking keeps the words that remain as they are and
// growing adds zero words, so a bit past size() can reappear after a shrink
// and a grow; ~ and flip() invert whole words; == compares the words (the
// reference's operator== is std::vector<uint64_t>'s).
#pragma once

#include <compare>
#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <iterator>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace seqan::hibf
{

namespace detail
{

// An allocator whose storage starts on a 64-byte boundary (one cache line).
template <typename T>
struct cache_line_allocator
{
    using value_type = T;
    static constexpr std::size_t alignment = 64u;

    cache_line_allocator() noexcept = default;
    template <typename U>
    constexpr cache_line_allocator(cache_line_allocator<U> const &) noexcept
    {}

    T * allocate(std::size_t const n)
    {
        return static_cast<T *>(::operator new(n * sizeof(T), std::align_val_t{alignment}));
    }
    void deallocate(T * const p, std::size_t) noexcept
    {
        ::operator delete(p, std::align_val_t{alignment});
    }

    template <typename U>
    bool operator==(cache_line_allocator<U> const &) const noexcept
    {
        return true;
    }
};

} // namespace detail

class bit_vector
{
    using word_t = uint64_t;
    using words_t = std::vector<word_t, detail::cache_line_allocator<word_t>>;

    static constexpr size_t words_for(size_t const bits) noexcept
    {
        return (bits + 63u) >> 6;
    }

    template <bool is_const>
    class bit_iterator;

    template <bool is_const>
    class bit_reference
    {
        using word_ptr = std::conditional_t<is_const, word_t const *, word_t *>;

        word_ptr word_;
        word_t mask_;

        constexpr bit_reference(word_ptr const word, size_t const bit) noexcept :
            word_{word},
            mask_{word_t{1} << (bit & 63u)}
        {}

        friend class bit_vector;
        template <bool>
        friend class bit_iterator;

    public:
        bit_reference() = delete;
        bit_reference(bit_reference const &) = default;
        bit_reference(bit_reference &&) = default;

        // Assignment from another proxy copies the bit it refers to, as for std::vector<bool>::reference.
        bit_reference & operator=(bit_reference const & other) noexcept
            requires (!is_const)
        {
            return *this = static_cast<bool>(other);
        }
        bit_reference & operator=(bit_reference && other) noexcept
            requires (!is_const)
        {
            return *this = static_cast<bool>(other);
        }
        constexpr bit_reference & operator=(bool const bit) noexcept
            requires (!is_const)
        {
            write(bit);
            return *this;
        }
        // Assignment through a const proxy: what makes `iterator` an output iterator of bool.
        constexpr bit_reference const & operator=(bool const bit) const noexcept
            requires (!is_const)
        {
            write(bit);
            return *this;
        }

        constexpr operator bool() const noexcept
        {
            return (*word_ & mask_) != 0u;
        }

        constexpr bit_reference & flip() noexcept
            requires (!is_const)
        {
            *word_ ^= mask_;
            return *this;
        }

    private:
        constexpr void write(bool const bit) const noexcept
        {
            *word_ = bit ? (*word_ | mask_) : (*word_ & ~mask_);
        }
    };

    // A position in the bits: the first word and a bit index counted from it.
    template <bool is_const>
    class bit_iterator
    {
        using word_ptr = std::conditional_t<is_const, word_t const *, word_t *>;

        word_ptr words_{};
        std::ptrdiff_t bit_{};

        template <bool>
        friend class bit_iterator;
        friend class bit_vector;

        constexpr bit_iterator(word_ptr const words, std::ptrdiff_t const bit) noexcept : words_{words}, bit_{bit} {}

    public:
        using value_type = bool;
        using reference = bit_reference<is_const>;
        using pointer = void;
        using difference_type = std::ptrdiff_t;
        using iterator_category = std::random_access_iterator_tag;
        using iterator_concept = std::random_access_iterator_tag;

        bit_iterator() = default;
        constexpr bit_iterator(bit_iterator<!is_const> const & other) noexcept
            requires is_const
            : words_{other.words_}, bit_{other.bit_}
        {}

        constexpr reference operator*() const noexcept
        {
            return reference{words_ + (bit_ >> 6), static_cast<size_t>(bit_)};
        }
        constexpr reference operator[](difference_type const n) const noexcept
        {
            return *(*this + n);
        }

        constexpr bit_iterator & operator++() noexcept
        {
            ++bit_;
            return *this;
        }
        constexpr bit_iterator operator++(int) noexcept
        {
            bit_iterator const old{*this};
            ++bit_;
            return old;
        }
        constexpr bit_iterator & operator--() noexcept
        {
            --bit_;
            return *this;
        }
        constexpr bit_iterator operator--(int) noexcept
        {
            bit_iterator const old{*this};
            --bit_;
            return old;
        }
        constexpr bit_iterator & operator+=(difference_type const n) noexcept
        {
            bit_ += n;
            return *this;
        }
        constexpr bit_iterator & operator-=(difference_type const n) noexcept
        {
            bit_ -= n;
            return *this;
        }
        constexpr bit_iterator operator+(difference_type const n) const noexcept
        {
            return bit_iterator{words_, bit_ + n};
        }
        friend constexpr bit_iterator operator+(difference_type const n, bit_iterator const it) noexcept
        {
            return it + n;
        }
        constexpr bit_iterator operator-(difference_type const n) const noexcept
        {
            return bit_iterator{words_, bit_ - n};
        }

        // Iterators into one bit_vector share words_: positions compare as bit indices.
        template <bool other_const>
        constexpr difference_type operator-(bit_iterator<other_const> const & other) const noexcept
        {
            return bit_ - other.bit_;
        }
        template <bool other_const>
        constexpr bool operator==(bit_iterator<other_const> const & other) const noexcept
        {
            return words_ == other.words_ && bit_ == other.bit_;
        }
        template <bool other_const>
        constexpr std::strong_ordering operator<=>(bit_iterator<other_const> const & other) const noexcept
        {
            return bit_ <=> other.bit_;
        }
    };

public:
    using value_type = bool;
    using reference = bit_reference<false>;
    using const_reference = bit_reference<true>;
    using iterator = bit_iterator<false>;
    using const_iterator = bit_iterator<true>;
    using size_type = size_t;
    using difference_type = std::ptrdiff_t;
    using allocator_type = detail::cache_line_allocator<bool>;

    bit_vector() = default;
    bit_vector(bit_vector const &) = default;
    bit_vector(bit_vector &&) = default;
    bit_vector & operator=(bit_vector const &) = default;
    bit_vector & operator=(bit_vector &&) = default;
    ~bit_vector() = default;

    // count bits of value bit; the words are filled whole.
    bit_vector(size_type const count, bool const bit)
    {
        assign(count, bit);
    }
    // count zero bits.
    bit_vector(size_type const count) : bit_vector(count, false)
    {}
    bit_vector(std::initializer_list<bool> const bits)
    {
        assign(bits);
    }

    void assign(size_type const count, bool const bit)
    {
        resize(count);
        for (word_t & w : words_)
            w = bit ? ~word_t{0} : word_t{0};
    }
    void assign(std::initializer_list<bool> const bits)
    {
        assign(bits.begin(), bits.end());
    }
    template <std::input_iterator it_t, std::sentinel_for<it_t> sentinel_t>
    void assign(it_t first, sentinel_t const last)
    {
        bit_vector tmp{};
        if constexpr (std::sized_sentinel_for<sentinel_t, it_t>)
            tmp.reserve(static_cast<size_type>(last - first));
        for (; first != last; ++first)
            tmp.push_back(static_cast<bool>(*first));
        swap(tmp);
    }

    // --- element access ----------------------------------------------------
    reference operator[](size_type const i) noexcept
    {
        return reference{words_.data() + (i >> 6), i};
    }
    const_reference operator[](size_type const i) const noexcept
    {
        return const_reference{words_.data() + (i >> 6), i};
    }
    reference back() noexcept
    {
        return (*this)[size_ - 1u];
    }
    const_reference back() const noexcept
    {
        return (*this)[size_ - 1u];
    }

    bool all() const noexcept
    {
        for (word_t const w : words_)
            if (w != ~word_t{0})
                return false;
        return true;
    }
    bool any() const noexcept
    {
        for (word_t const w : words_)
            if (w != 0u)
                return true;
        return false;
    }
    bool none() const noexcept
    {
        return !any();
    }

    // --- capacity ----------------------------------------------------------
    size_type size() const noexcept
    {
        return size_;
    }
    bool empty() const noexcept
    {
        return size_ == 0u;
    }
    size_type capacity() const noexcept
    {
        return words_.capacity() * 64u;
    }
    void reserve(size_type const bits)
    {
        words_.reserve(words_for(bits));
    }

    // --- modifiers ---------------------------------------------------------
    void push_back(bool const bit)
    {
        resize(size_ + 1u);
        back() = bit;
    }
    // Growing sets the new bits to bit (new words start at zero); shrinking drops the words past the new end.
    void resize(size_type const count, bool const bit = false)
    {
        words_.resize(words_for(count));
        size_type const old = size_;
        size_ = count;
        if (bit)
            for (size_type i = old; i < count; ++i)
                (*this)[i] = true;
    }
    // size() becomes 0; capacity() stays.
    void clear() noexcept
    {
        words_.clear();
        size_ = 0u;
    }
    void swap(bit_vector & other) noexcept
    {
        words_.swap(other.words_);
        std::swap(size_, other.size_);
    }

    // The word-wise operators take a right-hand side of the same size.
    bit_vector & operator&=(bit_vector const & rhs) noexcept
    {
        return combine(rhs, [](word_t const a, word_t const b) { return a & b; });
    }
    bit_vector & operator|=(bit_vector const & rhs) noexcept
    {
        return combine(rhs, [](word_t const a, word_t const b) { return a | b; });
    }
    bit_vector & operator^=(bit_vector const & rhs) noexcept
    {
        return combine(rhs, [](word_t const a, word_t const b) { return a ^ b; });
    }
    // this &= ~rhs without a temporary.
    bit_vector & and_not(bit_vector const & rhs) noexcept
    {
        return combine(rhs, [](word_t const a, word_t const b) { return a & ~b; });
    }
    bit_vector operator~() const
    {
        bit_vector out{*this};
        return out.flip();
    }
    friend bit_vector operator&(bit_vector lhs, bit_vector const & rhs) noexcept
    {
        return lhs &= rhs;
    }
    friend bit_vector operator|(bit_vector lhs, bit_vector const & rhs) noexcept
    {
        return lhs |= rhs;
    }
    friend bit_vector operator^(bit_vector lhs, bit_vector const & rhs) noexcept
    {
        return lhs ^= rhs;
    }
    bit_vector & flip() noexcept
    {
        for (word_t & w : words_)
            w = ~w;
        return *this;
    }
    bit_vector & flip(size_type const position)
    {
        if (position >= size_)
            throw std::out_of_range{"bit_vector::flip: position " + std::to_string(position)
                                    + " is out of the range [0, " + std::to_string(size_) + ")"};
        (*this)[position].flip();
        return *this;
    }

    friend bool operator==(bit_vector const & a, bit_vector const & b) noexcept
    {
        return a.words_ == b.words_;
    }

    // --- iterators ---------------------------------------------------------
    iterator begin() noexcept
    {
        return iterator{words_.data(), 0};
    }
    const_iterator begin() const noexcept
    {
        return const_iterator{words_.data(), 0};
    }
    const_iterator cbegin() const noexcept
    {
        return begin();
    }
    iterator end() noexcept
    {
        return iterator{words_.data(), static_cast<difference_type>(size_)};
    }
    const_iterator end() const noexcept
    {
        return const_iterator{words_.data(), static_cast<difference_type>(size_)};
    }
    const_iterator cend() const noexcept
    {
        return end();
    }

    // --- the words ---------------------------------------------------------
    word_t * data() noexcept
    {
        return words_.data();
    }
    word_t const * data() const noexcept
    {
        return words_.data();
    }
    // ceil(size() / 64)
    size_type word_count() const noexcept
    {
        return words_.size();
    }

private:
    template <typename op_t>
    bit_vector & combine(bit_vector const & rhs, op_t const op) noexcept
    {
        word_t * const a = words_.data();
        word_t const * const b = rhs.words_.data();
        size_t const n = words_.size() < rhs.words_.size() ? words_.size() : rhs.words_.size();
        for (size_t i = 0; i < n; ++i)
            a[i] = op(a[i], b[i]);
        return *this;
    }

    words_t words_{};
    size_type size_{};
};

} // namespace seqan::hibf
