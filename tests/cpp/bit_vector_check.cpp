// Checks of seqan::hibf::bit_vector (include/hibf/misc/bit_vector.hpp) on the
// host, printing one "<name> ok" line per group and exiting non-zero at the
// first failure:
//   - the expectations of the reference's test/unit/hibf/bit_vector_test.cpp
//     (construction :57-138, assign :140-213, iterators :215-299, element
//     access :301-320, all/any/none :322-384, resize/clear/push_back/swap
//     :386-458, &, |, ^, ~ :460-651, flip :653-742, size/empty :744-802,
//     output iterator :804-815), restated;
//   - the iterator concepts the reference's iterator_fixture and
//     output_iterator tests require;
//   - random sequences of resize, push_back, assignment through the proxy,
//     flip, the word-wise operators and swap against a model that keeps the
//     bits of whole 64-bit words in a std::vector<bool> (the reference's
//     word-wise semantics: bits past size() in the last word stay as they
//     are).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <iterator>
#include <random>
#include <ranges>
#include <sstream>
#include <stdexcept>
#include <utility>
#include <vector>

#include <hibf/misc/bit_vector.hpp>
#include <hibf/misc/print.hpp>

using seqan::hibf::bit_vector;

static_assert(std::random_access_iterator<bit_vector::iterator>);
static_assert(std::random_access_iterator<bit_vector::const_iterator>);
static_assert(std::output_iterator<bit_vector::iterator, bool>);
static_assert(!std::output_iterator<bit_vector::const_iterator, bool>);
static_assert(std::ranges::random_access_range<bit_vector>);
static_assert(std::ranges::sized_range<bit_vector>);
static_assert(std::same_as<bit_vector::value_type, bool>);
static_assert(std::same_as<bit_vector::size_type, size_t>);
static_assert(std::same_as<bit_vector::difference_type, std::ptrdiff_t>);
static_assert(std::same_as<bit_vector::reference, std::iter_reference_t<bit_vector::iterator>>);
static_assert(std::same_as<bit_vector::const_reference, std::iter_reference_t<bit_vector::const_iterator>>);
static_assert(std::convertible_to<bit_vector::iterator, bit_vector::const_iterator>);
static_assert(!std::is_assignable_v<bit_vector::const_reference, bool>);

namespace
{

int failures = 0;

void expect(bool const ok, char const * what, int line)
{
    if (!ok)
    {
        std::fprintf(stderr, "line %d: %s\n", line, what);
        ++failures;
    }
}
#define EXPECT(c) expect((c), #c, __LINE__)

void done(char const * name)
{
    if (failures)
        std::exit(1);
    std::printf("%s ok\n", name);
}

// Bits 0, 10, 63, 64, 127, 128, 200 and 249 of a 250-bit vector: either side of each word boundary.
constexpr size_t marks[] = {0, 10, 63, 64, 127, 128, 200, 249};

bool marked(size_t const i)
{
    return std::ranges::find(marks, i) != std::end(marks);
}

void construction_and_assignment()
{
    for (size_t n : {1000u, 64u, 512u, 1u})
    {
        bit_vector v(n);
        EXPECT(v.size() == n && v.none());
        bit_vector copy{v};
        bit_vector moved{std::move(copy)};
        bit_vector a{};
        a = moved;
        bit_vector b{};
        b = std::move(moved);
        EXPECT(a.size() == n && b.size() == n);
    }
    bit_vector ones(1000, true);
    EXPECT(ones.size() == 1000u);
    EXPECT(std::ranges::all_of(ones, [](bool b) { return b; }));

    bit_vector listed{true, false, true, false, false, true, true};
    bit_vector from_ints{1, 0, 1, 0, 0, 1, 1};
    bool const want[] = {true, false, true, false, false, true, true};
    EXPECT(listed.size() == 7u && from_ints.size() == 7u);
    for (size_t i = 0; i < 7; ++i)
        EXPECT(listed[i] == want[i] && from_ints[i] == want[i]);

    for (auto [n, bit] : {std::pair{1000u, false}, std::pair{64u, true}, std::pair{0u, false}, std::pair{1u, true}})
    {
        bit_vector v{};
        v.assign(n, bit);
        EXPECT(v.size() == n);
        EXPECT(std::ranges::all_of(v, [bit](bool b) { return b == bit; }));
    }
    bit_vector v{};
    v.assign({true, false, true, false, false, true, true});
    EXPECT(v == listed && v.size() == 7u);
    v.assign({1, 0, 1, 0, 0, 1, 1});
    EXPECT(v == listed);
    v.assign(std::initializer_list<bool>{});
    EXPECT(v.size() == 0u && v.empty());
    std::vector<bool> const src{false, true, true};
    v.assign(src.begin(), src.end());
    EXPECT(v.size() == 3u && !v[0] && v[1] && v[2]);
    done("construction_and_assignment");
}

void iterators_and_access()
{
    for (size_t n : {1000u, 64u})
    {
        bit_vector v(n, true);
        EXPECT(*v.begin());
        EXPECT(*std::as_const(v).begin() && *std::ranges::cbegin(std::as_const(v)));
        EXPECT(v.begin() != v.end() && std::as_const(v).begin() != std::ranges::cend(std::as_const(v)));
        EXPECT(v.end() - v.begin() == static_cast<std::ptrdiff_t>(n));
    }
    bit_vector empty{};
    EXPECT(empty.begin() == empty.end() && std::as_const(empty).begin() == std::as_const(empty).end());

    bit_vector seventy(70, true);
    auto it = seventy.begin();
    for (size_t i = 0; i < seventy.size(); ++i, ++it)
        EXPECT(*it);
    EXPECT(it == seventy.end());

    bit_vector v{true, false, true, false, false, true, true};
    EXPECT(v[0] && !v[1] && v[2]);
    EXPECT(!std::as_const(v)[3] && !std::as_const(v)[4] && std::as_const(v)[5] && std::as_const(v)[6]);
    EXPECT(v.back() && std::as_const(v).back());

    // random access both ways, across word boundaries, mixed with const iterators
    bit_vector r(300);
    for (size_t i = 0; i < 300; i += 3)
        r[i] = true;
    auto a = r.begin() + 130;
    EXPECT(*a == (130 % 3 == 0) && a[-1] == (129 % 3 == 0) && a[65] == (195 % 3 == 0));
    a -= 70;
    EXPECT(a - r.begin() == 60 && *a);
    auto b = 64 + a;
    EXPECT(b - a == 64 && a < b && b > a && a <= a && *b == (124 % 3 == 0));
    bit_vector::const_iterator c = b;
    EXPECT(c == b && b - c == 0 && !(c < b));
    --b;
    b--;
    EXPECT(b - a == 62);
    EXPECT((std::as_const(r).end() - r.begin()) == 300);
    done("iterators_and_access");
}

void all_any_none()
{
    EXPECT(bit_vector{}.all() && !bit_vector{}.any() && bit_vector{}.none());
    bit_vector ones(250, true);
    EXPECT(ones.all());
    ones[249] = false;
    EXPECT(!ones.all());
    ones[249] = true;
    ones[0] = false;
    EXPECT(!ones.all());

    bit_vector zeros(250, false);
    EXPECT(!zeros.any() && zeros.none());
    zeros[249] = true;
    EXPECT(zeros.any() && !zeros.none());
    zeros[249] = false;
    zeros[0] = true;
    EXPECT(zeros.any() && !zeros.none());
    done("all_any_none");
}

void modifiers()
{
    bit_vector v{};
    EXPECT(v.size() == 0u);
    v.resize(64);
    EXPECT(v.size() == 64u && v.none());
    v.resize(128, true);
    EXPECT(v.size() == 128u && v.any());
    v.resize(1, true); // the first word holds no set bit
    EXPECT(v.size() == 1u && v.none());
    v.resize(0, true);
    EXPECT(v.size() == 0u && v.none());

    bit_vector c{};
    size_t cap = c.capacity();
    c.clear();
    EXPECT(c.size() == 0u && c.capacity() == cap);
    c.resize(128, true);
    cap = c.capacity();
    EXPECT(c.size() == 128u && c.capacity() >= 128u);
    c.clear();
    EXPECT(c.size() == 0u && c.capacity() == cap);
    c.reserve(1000);
    EXPECT(c.capacity() >= 1000u && c.size() == 0u);

    bit_vector p{};
    p.push_back(true);
    EXPECT(p.size() == 1u && p.back());
    p.resize(128, true);
    p.push_back(false);
    EXPECT(p.size() == 129u && !p.back());

    bit_vector left{};
    bit_vector right(250, true);
    left.swap(right);
    EXPECT(left.size() == 250u && right.size() == 0u && left.all());
    right.resize(78);
    left.swap(right);
    EXPECT(left.size() == 78u && right.size() == 250u && left.none() && right.all());
    done("modifiers");
}

void word_wise_operators()
{
    bit_vector const all_set(250, true);
    {
        bit_vector v(250, false);
        v &= all_set;
        EXPECT(v.size() == 250u && v.none() && all_set.all());
        for (size_t i : marks)
            v[i] = true;
        v = v & all_set;
        EXPECT(!v.none());
        for (size_t i = 0; i < 250; ++i)
            EXPECT(v[i] == marked(i));
    }
    // | and ^ with bits 0, 10, 127, 128 and 249 cleared on the right as well
    auto cleared = [](bit_vector v, std::initializer_list<size_t> at)
    {
        for (size_t i : at)
            v[i] = false;
        return v;
    };
    bit_vector const rhs = cleared(all_set, {0, 10, 127, 128, 249});
    {
        bit_vector v(250, false);
        v |= all_set;
        EXPECT(v.all());
        v = cleared(v, {0, 10, 63, 64, 127, 128, 200, 249}) | rhs;
        EXPECT(!v.all() && !rhs.all());
        for (size_t i = 0; i < 250; ++i)
            EXPECT(v[i] == !(i == 0 || i == 10 || i == 127 || i == 128 || i == 249));
    }
    {
        bit_vector v(250, false);
        v ^= all_set;
        EXPECT(v.all());
        v = cleared(v, {0, 10, 63, 64, 127, 128, 200, 249}) ^ rhs;
        EXPECT(!v.all());
        for (size_t i = 0; i < 250; ++i)
            EXPECT(v[i] == (i == 63 || i == 64 || i == 200));
    }
    {
        bit_vector v(250, false);
        bit_vector n = ~v;
        EXPECT(n.size() == 250u && v.none() && n.all());
        for (size_t i : marks)
            v[i] = true;
        n = ~v;
        EXPECT(!n.all());
        for (size_t i = 0; i < 250; ++i)
            EXPECT(n[i] == !marked(i));
    }
    {
        bit_vector v(250, true);
        bit_vector m(250, false);
        for (size_t i : marks)
            m[i] = true;
        v.and_not(m);
        for (size_t i = 0; i < 250; ++i)
            EXPECT(v[i] == !marked(i));
    }
    done("word_wise_operators");
}

void flips()
{
    bit_vector v(250, false);
    v.flip();
    EXPECT(v.all());
    for (size_t i : marks)
        v[i] = false;
    v.flip();
    for (size_t i = 0; i < 250; ++i)
        EXPECT(v[i] == marked(i));

    bit_vector s(250, false);
    for (size_t i : marks)
        s.flip(i);
    for (size_t i = 0; i < 250; ++i)
        EXPECT(s[i] == marked(i));
    for (size_t i : {0u, 10u, 63u, 64u})
        s.flip(i);
    EXPECT(!s[0] && !s[10] && !s[63] && !s[64] && s[127]);
    bool threw = false;
    try
    {
        s.flip(250);
    }
    catch (std::out_of_range const &)
    {
        threw = true;
    }
    EXPECT(threw);
    s[5].flip();
    EXPECT(s[5]);
    done("flips");
}

void size_empty_output()
{
    EXPECT(bit_vector(1000).size() == 1000u && bit_vector(64).size() == 64u && bit_vector(1).size() == 1u);
    EXPECT(bit_vector{}.size() == 0u && bit_vector{}.empty() && !bit_vector(1).empty() && !bit_vector(1000).empty());

    bit_vector v(100, true);
    for (auto it = v.begin(); it != v.end(); it += 2)
        *it = false;
    for (auto it = v.begin(); it != v.end(); ++it)
        EXPECT(*it == ((it - v.begin()) % 2 == 1));
    // proxies assign bits, not positions
    v[0] = v[1];
    EXPECT(v[0] && v[1]);
    auto r = v[2];
    r = v[3];
    EXPECT(v[2]);
    std::ranges::fill(v, true);
    EXPECT(v.all());
    std::ranges::copy(std::vector<bool>(100, false), v.begin());
    EXPECT(std::ranges::none_of(v, [](bool b) { return b; }));
    EXPECT(v.any()); // bits 100..127 of the last word were set by the constructor and stay set

    // print formats a bit_vector as 0/1 (src/misc/print.cpp prints bool as uint16_t), also after a
    // caller set boolalpha on the stream, as the reference's snippets do before printing agent results.
    bit_vector p(5);
    p[1] = true;
    p[4] = true;
    std::ostringstream os;
    os << std::boolalpha;
    seqan::hibf::print(p, os);
    seqan::hibf::print(bit_vector{}, os);
    EXPECT(os.str() == "[0,1,0,0,1]\n[]\n");
    done("size_empty_output");
}

void data_layout()
{
    bit_vector v(200);
    v[0] = true;
    v[65] = true;
    v[199] = true;
    EXPECT(v.word_count() == 4u);
    EXPECT(v.data()[0] == 1u && v.data()[1] == 2u && v.data()[2] == 0u && v.data()[3] == (uint64_t{1} << 7));
    EXPECT(reinterpret_cast<uintptr_t>(v.data()) % 64u == 0u);
    done("data_layout");
}

// The model: the bits of all words, as the reference's word-wise operations see them.
struct model
{
    std::vector<bool> words_bits; // 64 * words
    size_t size{};

    void resize(size_t n, bool bit)
    {
        size_t const words = (n + 63) / 64;
        words_bits.resize(words * 64, false);
        size_t const old = size;
        size = n;
        if (bit)
            for (size_t i = old; i < n; ++i)
                words_bits[i] = true;
    }
};

bool same(bit_vector const & v, model const & m)
{
    if (v.size() != m.size || v.word_count() * 64 != m.words_bits.size())
        return false;
    for (size_t i = 0; i < m.words_bits.size(); ++i)
        if (((v.data()[i / 64] >> (i % 64)) & 1u) != m.words_bits[i])
            return false;
    bool all = true, any = false;
    for (bool b : m.words_bits)
        all &= b, any |= b;
    return v.all() == all && v.any() == any && v.none() == !any;
}

void random_against_model()
{
    std::mt19937_64 rng{7};
    for (int trial = 0; trial < 200; ++trial)
    {
        bit_vector a{}, b{};
        model ma{}, mb{};
        for (int step = 0; step < 300; ++step)
        {
            unsigned const op = rng() % 9;
            if (op == 0)
            {
                size_t const n = rng() % 400;
                bool const bit = rng() & 1u;
                a.resize(n, bit);
                ma.resize(n, bit);
            }
            else if (op == 1)
            {
                bool const bit = rng() & 1u;
                a.push_back(bit);
                ma.resize(ma.size + 1, false);
                ma.words_bits[ma.size - 1] = bit;
            }
            else if (op == 2 && a.size())
            {
                size_t const i = rng() % a.size();
                bool const bit = rng() & 1u;
                a[i] = bit;
                ma.words_bits[i] = bit;
            }
            else if (op == 3 && a.size())
            {
                size_t const i = rng() % a.size();
                a.flip(i);
                ma.words_bits[i] = !ma.words_bits[i];
            }
            else if (op == 4)
            {
                a.flip();
                ma.words_bits.flip();
            }
            else if (op >= 5 && op <= 7)
            {
                // a second vector of the same size, then a word-wise operator
                b = bit_vector(a.size());
                mb = model{};
                mb.resize(a.size(), false);
                for (size_t i = 0; i < b.word_count() * 64; ++i)
                    if (rng() & 1u)
                    {
                        b.data()[i / 64] |= uint64_t{1} << (i % 64);
                        mb.words_bits[i] = true;
                    }
                for (size_t i = 0; i < ma.words_bits.size(); ++i)
                {
                    bool const x = ma.words_bits[i], y = mb.words_bits[i];
                    ma.words_bits[i] = op == 5 ? (x && y) : op == 6 ? (x || y) : (x != y);
                }
                a = op == 5 ? (a & b) : op == 6 ? (a | b) : (a ^ b);
            }
            else if (op == 8)
            {
                bit_vector c = ~a;
                EXPECT(c.size() == a.size());
                c.flip();
                EXPECT(c == a);
                a.swap(b);
                std::swap(ma, mb);
            }
            if (!same(a, ma))
            {
                std::fprintf(stderr, "trial %d step %d op %u: differs from the model\n", trial, step, op);
                std::exit(1);
            }
        }
    }
    done("random_against_model");
}

} // namespace

int main()
{
    construction_and_assignment();
    iterators_and_access();
    all_any_none();
    modifiers();
    word_wise_operators();
    flips();
    size_empty_output();
    data_layout();
    random_against_model();
    return failures ? 1 : 0;
}
