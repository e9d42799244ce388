// The reference's snippet programs (test/snippet/ibf/*.cpp, hibf/*.cpp,
// readme.cpp) written against this repository's seqan::hibf headers.  Each
// section prints what the reference snippet prints, so the output can be
// compared with the reference's committed .out files.
#include <cstdint>
#include <iostream>
#include <ranges>
#include <vector>

#include <hibf/config.hpp>
#include <hibf/hierarchical_interleaved_bloom_filter.hpp>
#include <hibf/interleaved_bloom_filter.hpp>
#include <hibf/misc/counting_vector.hpp>
#include <hibf/misc/print.hpp>

namespace sh = seqan::hibf;

static void counting_agent()
{
    std::cout << "== ibf/counting_agent\n";
    sh::interleaved_bloom_filter ibf{sh::bin_count{8u}, sh::bin_size{8192u}, sh::hash_function_count{2u}};
    for (auto && value : std::views::iota(0u, 20u))
        ibf.emplace(value, sh::bin_index{0u});
    for (auto && value : std::views::iota(10u, 30u))
        ibf.emplace(value, sh::bin_index{4u});
    for (auto && value : std::views::iota(25u, 35u))
        ibf.emplace(value, sh::bin_index{7u});
    auto agent = ibf.counting_agent();
    auto & result = agent.bulk_count(std::views::iota(0u, 20u));
    sh::print(result);
    std::vector<size_t> const values{92, 1238, 812, 81273};
    sh::print(agent.bulk_count(values));
    sh::print(agent.bulk_count(std::views::iota(0u, 1024u)));
    auto agent2 = ibf.counting_agent<uint8_t>();
    sh::print(agent2.bulk_count(std::views::iota(0u, 20u)));
}

static void clear()
{
    std::cout << "== ibf/interleaved_bloom_filter_clear\n";
    sh::interleaved_bloom_filter ibf{sh::bin_count{8u}, sh::bin_size{8192u}, sh::hash_function_count{2u}};
    ibf.emplace_range(std::views::iota(0u, 20u), sh::bin_index{0u});
    ibf.emplace_range(std::views::iota(10u, 30u), sh::bin_index{4u});
    ibf.emplace_range(std::views::iota(25u, 35u), sh::bin_index{7u});
    auto agent = ibf.counting_agent();
    sh::print(agent.bulk_count(std::views::iota(0u, 20u)));
    ibf.clear(sh::bin_index{0u});
    sh::print(agent.bulk_count(std::views::iota(0u, 20u)));
    sh::print(agent.bulk_count(std::views::iota(0u, 1024u)));
    ibf.clear(std::vector{sh::bin_index{4u}, sh::bin_index{7u}});
    sh::print(agent.bulk_count(std::views::iota(0u, 1024u)));
}

static void containment()
{
    std::cout << "== ibf/containment_agent_bulk_contains\n";
    sh::interleaved_bloom_filter ibf{sh::bin_count{12u}, sh::bin_size{8192u}};
    ibf.emplace(126, sh::bin_index{0u});
    ibf.emplace(712, sh::bin_index{3u});
    ibf.emplace(237, sh::bin_index{9u});
    auto agent = ibf.containment_agent();
    auto & result = agent.bulk_contains(712);
    sh::print(result);
    ibf.increase_bin_number_to(sh::bin_count{60u});
    agent = ibf.containment_agent();
}

static void counting_vector()
{
    std::cout << "== ibf/counting_vector\n";
    sh::interleaved_bloom_filter ibf{sh::bin_count{12u}, sh::bin_size{8192u}};
    ibf.emplace(126, sh::bin_index{0u});
    ibf.emplace(126, sh::bin_index{3u});
    ibf.emplace(126, sh::bin_index{9u});
    ibf.emplace(712, sh::bin_index{3u});
    ibf.emplace(237, sh::bin_index{9u});
    sh::counting_vector<uint16_t> counts(12, 0);
    auto agent = ibf.containment_agent();
    counts += agent.bulk_contains(712);
    sh::print(counts);
    counts += agent.bulk_contains(237);
    sh::print(counts);
    counts += agent.bulk_contains(126);
    sh::print(counts);
    counts += counts;
    sh::print(counts);
}

static void membership()
{
    std::cout << "== ibf/membership_agent_membership_for\n";
    sh::interleaved_bloom_filter ibf{sh::bin_count{12u}, sh::bin_size{8192u}};
    std::vector<uint64_t> const query{126, 712, 237};
    ibf.emplace(165, sh::bin_index{0u});
    for (auto && value : query)
    {
        ibf.emplace(value, sh::bin_index{3u});
        ibf.emplace(value, sh::bin_index{5u});
    }
    ibf.emplace(126, sh::bin_index{7u});
    ibf.emplace(712, sh::bin_index{7u});
    ibf.emplace(956, sh::bin_index{9u});
    auto agent = ibf.membership_agent();
    auto & result = agent.membership_for(query, 2u);
    sh::print(result);
}

static void increase_bin_number()
{
    std::cout << "== ibf/interleaved_bloom_filter_increase_bin_number_to\n";
    sh::interleaved_bloom_filter ibf{sh::bin_count{12u}, sh::bin_size{8192u}};
    ibf.emplace(126, sh::bin_index{0u});
    ibf.emplace(712, sh::bin_index{3u});
    ibf.emplace(237, sh::bin_index{9u});
    ibf.increase_bin_number_to(sh::bin_count{18u});
    auto agent = ibf.containment_agent();
    sh::print(agent.bulk_contains(126));
    sh::print(agent.bulk_contains(712));
    sh::print(agent.bulk_contains(237));
}

static void try_increase_bin_number()
{
    std::cout << "== ibf/interleaved_bloom_filter_try_increase_bin_number_to\n";
    sh::interleaved_bloom_filter ibf{sh::bin_count{73u}, sh::bin_size{1024u}};
    ibf.emplace(126, sh::bin_index{0u});
    ibf.emplace(712, sh::bin_index{3u});
    ibf.emplace(237, sh::bin_index{9u});
    for (size_t n : {73u, 50u, 128u, 129u})
    {
        bool const result = ibf.try_increase_bin_number_to(sh::bin_count{n});
        std::cout << std::boolalpha << result << '\n' << ibf.bin_count() << '\n';
    }
    auto agent = ibf.containment_agent();
    sh::print(agent.bulk_contains(126));
    sh::print(agent.bulk_contains(712));
    sh::print(agent.bulk_contains(237));
}

static void hibf()
{
    std::cout << "== hibf/hierarchical_interleaved_bloom_filter\n";
    std::vector<std::vector<size_t>> hashes{{1u, 2u, 3u, 4u, 5u, 6u, 7u, 8u, 9u, 10u}, {1u, 2u, 3u, 4u, 5u}};
    auto my_input = [&](size_t const user_bin_id, sh::insert_iterator it)
    {
        for (auto const hash : hashes[user_bin_id])
            it = hash;
    };
    sh::config config{.input_fn = my_input, .number_of_user_bins = 2};
    sh::hierarchical_interleaved_bloom_filter hibf{config};
    std::vector<size_t> query{1u, 2u, 3u};
    std::vector<size_t> query2{8u, 9u, 10u};
    auto agent = hibf.membership_agent();
    auto & result = agent.membership_for(query, 2u);
    sh::print(result);
    agent.sort_results();
    sh::print(result);
    auto & result2 = agent.membership_for(query2, 2u);
    sh::print(result2);
}

static void readme()
{
    std::cout << "== readme\n";
    std::vector<uint64_t> user_bin_1{1u, 2u, 3u, 4u, 5u, 6u, 7u, 8u, 9u, 10u};
    std::vector<uint64_t> user_bin_2{1u, 2u, 3u, 4u, 5u};
    std::vector<uint64_t> user_bin_3{3u, 9u, 11u};
    std::vector<std::vector<uint64_t>> user_bin_data{user_bin_1, user_bin_2, user_bin_3};
    auto get_user_bin_data = [&](size_t const user_bin_id, sh::insert_iterator it)
    {
        for (auto value : user_bin_data[user_bin_id])
            it = value;
    };
    sh::config config{.input_fn = get_user_bin_data,
                      .number_of_user_bins = 3u,
                      .number_of_hash_functions = 2u,
                      .maximum_fpr = 0.05,
                      .threads = 1u};
    sh::hierarchical_interleaved_bloom_filter hibf{config};
    std::vector<uint64_t> query1{3u, 9u, 12u, 14u};
    auto agent = hibf.membership_agent();
    auto & result1 = agent.membership_for(query1, 2u);
    for (uint64_t hit_user_bin : result1)
        std::cout << hit_user_bin << ' ';
    std::cout << '\n';
    auto query2 = std::views::iota(0u, 15u);
    auto & result2 = agent.membership_for(query2, 5u);
    agent.sort_results();
    for (uint64_t hit_user_bin : result2)
        std::cout << hit_user_bin << ' ';
    std::cout << '\n';
}

static void errors()
{
    std::cout << "== errors\n";
    auto expect_logic_error = [](auto make)
    {
        try
        {
            make();
            std::cout << "no error\n";
        }
        catch (std::logic_error const & e)
        {
            std::cout << e.what() << '\n';
        }
    };
    expect_logic_error([] { sh::interleaved_bloom_filter{sh::bin_count{64u}, sh::bin_size{0u}}; });
    expect_logic_error([] { sh::interleaved_bloom_filter{sh::bin_count{0u}, sh::bin_size{32u}}; });
    expect_logic_error([] { sh::interleaved_bloom_filter{sh::bin_count{64u}, sh::bin_size{32u}, sh::hash_function_count{0u}}; });
    expect_logic_error([] { sh::interleaved_bloom_filter{sh::bin_count{64u}, sh::bin_size{32u}, sh::hash_function_count{6u}}; });
    sh::config bad{.input_fn = [](size_t, sh::insert_iterator) {}, .number_of_user_bins = 0};
    expect_logic_error([&] { sh::hierarchical_interleaved_bloom_filter{bad}; });
}

int main()
{
    counting_agent();
    clear();
    containment();
    counting_vector();
    membership();
    increase_bin_number();
    try_increase_bin_number();
    hibf();
    readme();
    errors();
}
