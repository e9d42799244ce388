// seqan::hibf::layout::layout as a command-line filter for the tests: reads
// the layouts on stdin one after the other (a partitioned layout file) and,
// for each, prints its records ("top <id>", "max <i;j;...> <id>",
// "ub <idx> <prev;...> <storage> <count>") and then its write_to text.
#include <iostream>
#include <sstream>

#include <hibf/layout/layout.hpp>

static std::string joined(std::vector<size_t> const & v)
{
    std::string s;
    for (size_t i = 0; i < v.size(); ++i)
        s += (i ? ";" : "") + std::to_string(v[i]);
    return s.empty() ? "-" : s;
}

int main()
{
    while (std::cin.peek() != EOF)
    {
        seqan::hibf::layout::layout l;
        l.read_from(std::cin);
        std::cout << "top " << l.top_level_max_bin_id << '\n';
        for (auto const & m : l.max_bins)
            std::cout << "max " << joined(m.previous_TB_indices) << ' ' << m.id << '\n';
        for (auto const & u : l.user_bins)
            std::cout << "ub " << u.idx << ' ' << joined(u.previous_TB_indices) << ' ' << u.storage_TB_id << ' '
                      << u.number_of_technical_bins << '\n';
        std::ostringstream w;
        l.write_to(w);
        std::cout << w.str();
        seqan::hibf::layout::layout copy = l;
        copy.clear();
        std::cout << "cleared " << (copy.top_level_max_bin_id == 0 && copy.max_bins.empty() && copy.user_bins.empty())
                  << ' ' << (copy == l) << '\n';
    }
    return 0;
}
