// seqan::hibf::sketch::toolbox: the layout's use of HyperLogLog sketches
// (include/hibf/sketch/toolbox.hpp, src/sketch/toolbox.cpp in the reference).
//
// * sort_by_cardinalities orders user bins by decreasing k-mer count.
// * precompute_initial_union_estimates / precompute_union_estimates_for give
//   the DP of hierarchical_binning the estimated sizes of unions of contiguous
//   runs of the sorted user bins (prefix unions, and the unions ending at
//   column j).
// * rearrange_bins cuts the sorted order into intervals whose counts stay
//   within max_ratio of the interval's first bin, clusters each interval
//   agglomeratively by the Jaccard distance estimate
//   2 - (|A| + |B|) / |A u B| and lays the interval out in the tree's leaf
//   order, keeping one bin of overlap with the previous interval.
//
// Names, arguments and results are the reference's.  Two differences:
// the id map is a std::unordered_map (the reference uses
// robin_hood::unordered_flat_map; only lookups, inserts, erases and begin()
// on a one-element map are used, so the results do not depend on it), and
// cluster_bins runs its selection single-threaded whatever num_threads says:
// with several threads the reference picks among equally distant pairs in an
// order that depends on its dynamic OpenMP schedule, with one thread it is
// fixed, and this is that order.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <limits>
#include <queue>
#include <random>
#include <unordered_map>
#include <utility>
#include <vector>

#include <hibf/sketch/hyperloglog.hpp>

namespace seqan::hibf::sketch::toolbox
{

// toolbox.hpp:26-35: a node of the clustering tree; leaves have left == right == max.
struct clustering_node
{
    size_t left;
    size_t right;
    hyperloglog hll;
};

// toolbox.hpp:37-47
struct neighbor
{
    size_t id;
    double dist;

    bool operator>(neighbor const & other) const { return dist > other.dist; }
};

using prio_queue = std::priority_queue<neighbor, std::vector<neighbor>, std::greater<neighbor>>;

// toolbox.hpp:52-57
struct entry
{
    size_t id;
    prio_queue pq;
};

using distance_matrix = std::vector<entry>;
using id_map = std::unordered_map<size_t, size_t>;

// toolbox.cpp:23-33: positions by decreasing count (std::sort, as the reference).
inline void sort_by_cardinalities(std::vector<size_t> const & counts, std::vector<size_t> & positions)
{
    std::sort(positions.begin(),
              positions.end(),
              [&counts](size_t const a, size_t const b) { return counts[a] > counts[b]; });
}

// toolbox.cpp:35-51: estimates[j] = counts[positions[j]]; for j' < j, estimates[j'] is
// the estimate of the union of the sketches at positions[j' .. j].
inline void precompute_union_estimates_for(std::vector<uint64_t> & estimates,
                                           std::vector<hyperloglog> const & sketches,
                                           std::vector<size_t> const & counts,
                                           std::vector<size_t> const & positions,
                                           int64_t const j)
{
    hyperloglog running = sketches[positions[j]];
    estimates[j] = counts[positions[j]];
    for (int64_t k = j - 1; k >= 0; --k)
        estimates[k] = static_cast<uint64_t>(running.merge_and_estimate(sketches[positions[k]]));
}

// toolbox.cpp:53-70: estimates[0] = counts[positions[0]]; estimates[j] is the estimate
// of the union of the sketches at positions[0 .. j].
inline void precompute_initial_union_estimates(std::vector<uint64_t> & estimates,
                                               std::vector<hyperloglog> const & sketches,
                                               std::vector<size_t> const & counts,
                                               std::vector<size_t> const & positions)
{
    estimates.resize(sketches.size());
    hyperloglog running = sketches[positions[0]];
    estimates[0] = counts[positions[0]];
    for (size_t j = 1; j < positions.size(); ++j)
        estimates[j] = static_cast<uint64_t>(running.merge_and_estimate(sketches[positions[j]]));
}

// toolbox.cpp:385-399: Fisher-Yates over dist with std::mt19937_64 seeded
// 0x7E1E5665D46800E5 and std::uniform_int_distribution<size_t>(i, n - 1); the
// id map follows each moved entry.
inline void random_shuffle(distance_matrix & dist, id_map & remaining_ids)
{
    size_t const n = dist.size();
    std::mt19937_64 gen(0x7E1E5665D46800E5ULL);
    for (size_t i = 0; i + 1 < n; ++i)
    {
        std::uniform_int_distribution<size_t> pick(i, n - 1);
        size_t const other = pick(gen);
        size_t const id_i = dist[i].id;
        size_t const id_other = dist[other].id;
        std::swap(dist[i], dist[other]);
        std::swap(remaining_ids[id_i], remaining_ids[id_other]);
    }
}

// toolbox.cpp:401-438: moves the entries whose id is still in remaining_ids to the
// front (two cursors from both ends, swapping a dead front entry with a live back
// entry), updates their indices and drops the rest.
inline void prune(distance_matrix & dist, id_map & remaining_ids)
{
    if (dist.empty())
        return;
    auto live = [&remaining_ids](entry const & e) { return remaining_ids.count(e.id) != 0u; };
    size_t front = 0;
    size_t back = dist.size() - 1;
    while (front != back)
    {
        if (live(dist[front]))
            ++front;
        else if (!live(dist[back]))
            --back;
        else
        {
            std::swap(dist[front], dist[back]);
            remaining_ids.at(dist[front].id) = front;
        }
    }
    if (live(dist[front]))
        ++front;
    dist.resize(front);
}

// toolbox.cpp:440-470: swaps children on the path from `id` down to the leaf
// previous_rightmost so that it becomes the leftmost leaf.  Returns whether the
// subtree holds it.
inline bool rotate(std::vector<clustering_node> & clustering,
                   size_t const previous_rightmost,
                   size_t const first,
                   size_t const id)
{
    if (id == previous_rightmost)
        return true;
    clustering_node & node = clustering[id - first];
    if (node.left == std::numeric_limits<size_t>::max())
        return false;
    if (rotate(clustering, previous_rightmost, first, node.left))
        return true;
    if (rotate(clustering, previous_rightmost, first, node.right))
    {
        std::swap(node.left, node.right);
        return true;
    }
    return false;
}

// toolbox.cpp:472-490: leaves left to right, previous_rightmost left out.
inline void trace(std::vector<clustering_node> const & clustering,
                  std::vector<size_t> & permutation,
                  size_t const previous_rightmost,
                  size_t const first,
                  size_t const id)
{
    clustering_node const & node = clustering[id - first];
    if (node.left == std::numeric_limits<size_t>::max())
    {
        if (id != previous_rightmost)
            permutation.push_back(id);
        return;
    }
    trace(clustering, permutation, previous_rightmost, first, node.left);
    trace(clustering, permutation, previous_rightmost, first, node.right);
}

namespace detail
{

// Estimates of the unions of one sketch with each of a list of others, for
// the clustering: unions(a, others, out) sets out[k] = estimate(merge(a,
// *others[k])).  reset() is called when a clustering starts; within one, a
// sketch object is not modified once passed (so an implementation may cache
// it by address).  The default computes the estimates one by one on the host.
struct union_source
{
    virtual ~union_source() = default;
    virtual void reset() {}
    virtual void unions(hyperloglog const & a, std::vector<hyperloglog const *> const & others, double * out) = 0;
    // Every pair i < j of `leaves`: out[i L - i (i + 1) / 2 + (j - i - 1)] = estimate(merge(i, j)), L leaves.
    virtual void all_pairs(std::vector<hyperloglog const *> const & leaves, double * out)
    {
        size_t const n = leaves.size();
        std::vector<hyperloglog const *> others;
        for (size_t i = 0; i + 1 < n; ++i)
        {
            others.assign(leaves.begin() + i + 1, leaves.end());
            unions(*leaves[i], others, out + (i * n - i * (i + 1) / 2));
        }
    }
};

struct host_union_source final : union_source
{
    void unions(hyperloglog const & a, std::vector<hyperloglog const *> const & others, double * out) override
    {
        for (size_t k = 0; k < others.size(); ++k)
        {
            hyperloglog tmp = a;
            out[k] = tmp.merge_and_estimate(*others[k]);
        }
    }
};

} // namespace detail

// toolbox.cpp:109-383 with one thread.  Appends the ids first .. last - 1 to
// permutation in the leaf order of the clustering tree of
// sketches[positions[first .. last)] (plus, after the first interval, the last
// id of the previous one as an extra leaf that is rotated to the left and not
// appended).  `unions` computes the merged-sketch estimates.
inline void cluster_bins(std::vector<hyperloglog> const & sketches,
                         std::vector<size_t> & positions,
                         std::vector<size_t> & permutation,
                         size_t const first,
                         size_t const last,
                         size_t const /* num_threads */,
                         detail::union_source & unions)
{
    unions.reset();
    constexpr size_t none = std::numeric_limits<size_t>::max();
    size_t const prune_steps = static_cast<size_t>(std::floor(std::sqrt(sketches.size())));
    size_t steps_without_prune = 0;

    distance_matrix dist;
    id_map remaining_ids;
    std::vector<clustering_node> clustering;
    std::vector<double> estimates;
    clustering.reserve(2 * (last - first) + 2);
    estimates.reserve(2 * (last - first) + 2);

    size_t new_id = last - 1; // incremented before each use
    for (size_t id = first; id < last; ++id)
    {
        clustering.push_back({none, none, sketches[positions[id]]});
        estimates.push_back(clustering.back().hll.estimate());
    }
    size_t previous_rightmost = none;
    if (first != 0)
    {
        // The last leaf of the previous interval joins as id `last` (its index in
        // clustering); it only anchors the order and is not traced.
        size_t const anchor = permutation.back();
        previous_rightmost = ++new_id;
        clustering.push_back({none, none, sketches[positions[anchor]]});
        estimates.push_back(clustering.back().hll.estimate());
    }
    size_t const leaves = clustering.size();
    for (size_t k = 0; k < leaves; ++k)
    {
        dist.push_back({first + k, prio_queue{}});
        remaining_ids[first + k] = k;
    }

    // Distances of every leaf pair (i < j) into the queue of i; the first minimum
    // in leaf order seeds the selection.
    size_t min_id = none;
    double min_dist = std::numeric_limits<double>::max();
    {
        std::vector<hyperloglog const *> leaf_sketches(leaves);
        for (size_t i = 0; i < leaves; ++i)
            leaf_sketches[i] = &clustering[i].hll;
        std::vector<double> merged(leaves * (leaves - 1) / 2);
        if (leaves > 1)
            unions.all_pairs(leaf_sketches, merged.data());
        for (size_t i = 0; i < leaves; ++i)
        {
            double const * const row = merged.data() + (i * leaves - i * (i + 1) / 2);
            for (size_t j = i + 1; j < leaves; ++j)
                dist[i].pq.push({j + first, 2 - (estimates[i] + estimates[j]) / row[j - i - 1]});
            if (dist[i].pq.empty())
                continue;
            if (dist[i].pq.top().dist < min_dist)
            {
                min_dist = dist[i].pq.top().dist;
                min_id = dist[i].id;
            }
        }
    }

    random_shuffle(dist, remaining_ids);

    std::vector<hyperloglog const *> others;
    std::vector<size_t> other_index;
    std::vector<double> merged;
    while (remaining_ids.size() > 1)
    {
        ++new_id;
        size_t const min_index = remaining_ids.at(min_id);
        size_t const neighbor_id = dist[min_index].pq.top().id;

        // The new cluster takes over min_id's sketch and slot in dist.
        clustering.push_back({min_id, neighbor_id, std::move(clustering[min_id - first].hll)});
        estimates.push_back(clustering.back().hll.merge_and_estimate(clustering[neighbor_id - first].hll));
        remaining_ids.erase(min_id);
        remaining_ids.erase(neighbor_id);
        remaining_ids[new_id] = min_index;
        dist[min_index] = {new_id, prio_queue{}};
        if (++steps_without_prune > prune_steps)
        {
            prune(dist, remaining_ids);
            steps_without_prune = 0;
        }

        // Distance of the new cluster to every live one, pushed into that one's
        // queue; stale tops are dropped; the first minimum in dist order wins.
        hyperloglog const & fresh = clustering.back().hll;
        others.clear();
        other_index.clear();
        for (size_t i = 0; i < dist.size(); ++i)
        {
            size_t const other_id = dist[i].id;
            if (other_id == new_id || remaining_ids.count(other_id) == 0u)
                continue;
            others.push_back(&clustering[other_id - first].hll);
            other_index.push_back(i);
        }
        merged.resize(others.size());
        if (!others.empty())
            unions.unions(fresh, others, merged.data());
        min_id = none;
        min_dist = std::numeric_limits<double>::max();
        for (size_t k = 0; k < other_index.size(); ++k)
        {
            entry & e = dist[other_index[k]];
            e.pq.push({new_id, 2 - (estimates[e.id - first] + estimates.back()) / merged[k]});
            while (remaining_ids.count(e.pq.top().id) == 0u)
                e.pq.pop();
            if (e.pq.top().dist < min_dist)
            {
                min_dist = e.pq.top().dist;
                min_id = e.id;
            }
        }
    }

    size_t const root_id = dist[remaining_ids.begin()->second].id;
    if (first != 0)
        rotate(clustering, previous_rightmost, first, root_id);
    trace(clustering, permutation, previous_rightmost, first, root_id);
}

inline void cluster_bins(std::vector<hyperloglog> const & sketches,
                         std::vector<size_t> & positions,
                         std::vector<size_t> & permutation,
                         size_t const first,
                         size_t const last,
                         size_t const num_threads)
{
    detail::host_union_source host;
    cluster_bins(sketches, positions, permutation, first, last, num_threads, host);
}

// toolbox.cpp:72-107: intervals of the sorted order in which every count is at
// least max_ratio times the interval's first, clustered one after the other;
// the permutation is then applied to positions in place.
inline void rearrange_bins(std::vector<hyperloglog> const & sketches,
                           std::vector<size_t> const & kmer_counts,
                           std::vector<size_t> & positions,
                           double const max_ratio,
                           size_t const num_threads,
                           detail::union_source & unions)
{
    std::vector<size_t> permutation;
    size_t first = 0;
    for (size_t last = 1; first < positions.size(); ++last)
    {
        if (last == positions.size() || kmer_counts[positions[first]] * max_ratio > kmer_counts[positions[last]])
        {
            cluster_bins(sketches, positions, permutation, first, last, num_threads, unions);
            first = last;
        }
    }
    for (size_t i = 0; i < permutation.size(); ++i)
    {
        size_t target = permutation[i];
        while (target < i)
            target = permutation[target];
        std::swap(positions[i], positions[target]);
    }
}

inline void rearrange_bins(std::vector<hyperloglog> const & sketches,
                           std::vector<size_t> const & kmer_counts,
                           std::vector<size_t> & positions,
                           double const max_ratio,
                           size_t const num_threads)
{
    detail::host_union_source host;
    rearrange_bins(sketches, kmer_counts, positions, max_ratio, num_threads, host);
}

} // namespace seqan::hibf::sketch::toolbox
