// Known answers of the sketch toolbox, from test/unit/hibf/sketch/toolbox_test.cpp
// (libstdc++ expectations).  Prints one line per check and "ok" at the end;
// exits non-zero on the first mismatch.
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <random>
#include <vector>

#include <hibf/sketch/hyperloglog.hpp>
#include <hibf/sketch/toolbox.hpp>

namespace tb = seqan::hibf::sketch::toolbox;
using seqan::hibf::sketch::hyperloglog;

static int failures = 0;

template <typename A, typename B>
static void expect(char const * what, A const & got, B const & want)
{
    bool same = got.size() == want.size();
    for (size_t i = 0; same && i < got.size(); ++i)
        same = static_cast<unsigned long long>(got[i]) == static_cast<unsigned long long>(want[i]);
    std::printf("%s %s:", same ? "pass" : "FAIL", what);
    for (auto const v : got)
        std::printf(" %llu", static_cast<unsigned long long>(v));
    std::printf("\n");
    failures += !same;
}

int main()
{
    // toolbox_test.cpp:21-38: four copies of one 5-bit sketch of 1,500 values.
    std::vector<size_t> counts{500, 600, 700, 800};
    std::vector<size_t> positions{0, 1, 2, 3};
    std::vector<hyperloglog> sketches = [&]()
    {
        hyperloglog s{};
        std::uniform_int_distribution<uint64_t> dist{};
        std::mt19937_64 engine{0u};
        for (size_t i = 0; i < 1500; ++i)
            s.add(dist(engine));
        return std::vector<hyperloglog>(counts.size(), s);
    }();

    {   // :41-50
        std::vector<size_t> p = positions;
        tb::sort_by_cardinalities(counts, p);
        expect("sort_by_cardinalities", p, std::vector<size_t>{3, 2, 1, 0});
    }
    {   // :52-83
        std::vector<uint64_t> est(4);
        tb::precompute_union_estimates_for(est, sketches, counts, positions, 0);
        expect("union_estimates_for_0", est, std::vector<uint64_t>{500, 0, 0, 0});
        tb::precompute_union_estimates_for(est, sketches, counts, positions, 1);
        expect("union_estimates_for_1", est, std::vector<uint64_t>{1635, 600, 0, 0});
        tb::precompute_union_estimates_for(est, sketches, counts, positions, 2);
        expect("union_estimates_for_2", est, std::vector<uint64_t>{1635, 1635, 700, 0});
        tb::precompute_union_estimates_for(est, sketches, counts, positions, 3);
        expect("union_estimates_for_3", est, std::vector<uint64_t>{1635, 1635, 1635, 800});
    }
    {   // :85-115
        tb::distance_matrix dist{{0, {}}, {1, {}}, {2, {}}, {3, {}}, {4, {}}};
        tb::id_map ids{{0, 0}, {1, 1}, {2, 2}, {3, 3}, {4, 4}};
        tb::random_shuffle(dist, ids);
        expect("random_shuffle_ids", std::vector<size_t>{ids[0], ids[1], ids[2], ids[3], ids[4]},
               std::vector<size_t>{3, 2, 1, 0, 4});
        expect("random_shuffle_dist",
               std::vector<size_t>{dist[ids[0]].id, dist[ids[1]].id, dist[ids[2]].id, dist[ids[3]].id, dist[ids[4]].id},
               std::vector<size_t>{0, 1, 2, 3, 4});
    }
    {   // :117-158
        tb::distance_matrix dist{{0, {}}, {1, {}}, {2, {}}, {3, {}}, {4, {}}};
        tb::id_map ids{{0, 0}, {1, 1}, {2, 2}, {3, 3}, {4, 4}};
        tb::prune(dist, ids);
        expect("prune_all_live", std::vector<size_t>{dist.size(), dist[0].id, dist[1].id, dist[2].id, dist[3].id, dist[4].id},
               std::vector<size_t>{5, 0, 1, 2, 3, 4});
        ids.erase(1);
        ids.erase(3);
        tb::prune(dist, ids);
        expect("prune_ids", std::vector<size_t>{ids[0], ids[2], ids[4]}, std::vector<size_t>{0, 2, 1});
        expect("prune_dist", std::vector<size_t>{dist.size(), dist[0].id, dist[1].id, dist[2].id},
               std::vector<size_t>{3, 0, 4, 2});
    }
    size_t const f = std::numeric_limits<size_t>::max();
    {   // :160-203
        hyperloglog s{5};
        std::vector<tb::clustering_node> c{{f, f, s}, {f, f, s}, {f, f, s}, {f, f, s}, {5, 6, s}, {0, 1, s}, {2, 3, s}};
        tb::rotate(c, 0, 0, 4);
        expect("rotate_none", std::vector<size_t>{c[4].left, c[4].right, c[5].left, c[5].right, c[6].left, c[6].right},
               std::vector<size_t>{5, 6, 0, 1, 2, 3});
        tb::rotate(c, 2, 0, 4);
        expect("rotate_right", std::vector<size_t>{c[4].left, c[4].right, c[5].left, c[5].right, c[6].left, c[6].right},
               std::vector<size_t>{6, 5, 0, 1, 2, 3});
    }
    {   // :205-236
        hyperloglog s{5};
        std::vector<tb::clustering_node> c{{f, f, s}, {f, f, s}, {f, f, s}, {f, f, s}, {5, 6, s}, {1, 3, s}, {2, 0, s}};
        std::vector<size_t> perm;
        tb::trace(c, perm, 2, 0, 4);
        expect("trace", perm, std::vector<size_t>{1, 3, 0});
    }
    {   // :238-269
        std::vector<size_t> perm;
        std::vector<size_t> p = positions;
        tb::cluster_bins(sketches, p, perm, 0, 3, 1);
        expect("cluster_bins_whole", perm, std::vector<size_t>{2, 0, 1});
        perm.clear();
        tb::cluster_bins(sketches, p, perm, 0, 1, 1);
        expect("cluster_bins_first", perm, std::vector<size_t>{0});
        tb::cluster_bins(sketches, p, perm, 1, 3, 1);
        expect("cluster_bins_second", perm, std::vector<size_t>{0, 1, 2});
    }
    {   // :271-314
        std::vector<size_t> p = positions;
        std::vector<size_t> c = counts;
        tb::rearrange_bins(sketches, c, p, 0.9, 1);
        expect("rearrange_identical", p, std::vector<size_t>{0, 1, 2, 3});

        std::vector<size_t> kc(7, 5000);
        std::vector<size_t> pos{0, 1, 2, 3, 4, 5, 6};
        std::vector<hyperloglog> sk(kc.size(), hyperloglog(10));
        for (size_t i = 0; i < 5000; ++i)
        {
            sk[0].add(i);
            sk[3].add(i);
            sk[1].add(i + 4000);
            sk[6].add(i + 4000);
            sk[2].add(i + 20000);
            sk[4].add(i + 20000);
            sk[5].add(i + 8000);
        }
        tb::rearrange_bins(sk, kc, pos, 0.9, 1);
        expect("rearrange_overlaps", pos, std::vector<size_t>{5, 1, 6, 0, 3, 2, 4});
    }
    if (failures)
        return 1;
    std::printf("ok\n");
    return 0;
}
