// HIBF build plan from a layout: restates the reference's layout graph
// (src/layout/graph.cpp:19-125), construct_ibf (src/build/construct_ibf.cpp:26-64),
// insert_into_ibf's split-bin chunking (src/build/insert_into_ibf.cpp:28-69) and
// hierarchical_build (src/hierarchical_interleaved_bloom_filter.cpp:40-179).
// Instead of inserting values, it records which value ranges go to which
// technical bin; the device bulk emplace then sets the bits.
#include "build.hpp"

#include <algorithm>
#include <cmath>
#include <optional>
#include <string>

namespace hibf_amd::host
{

namespace
{

// Largest technical-bin count of one IBF a layout may ask for (far above any
// layout's; the device descent handles at most 8,192), so that a malformed
// layout cannot size tables of arbitrary length.
constexpr uint64_t kMaxLayoutBins = 1ULL << 24;

struct Node
{
    std::vector<Node> children;
    uint64_t parent_bin_index{};
    uint64_t max_bin_index{};
    uint64_t number_of_technical_bins{};
    std::optional<size_t> favourite_child_idx{};
    std::vector<Layout::UserBin> remaining_records{};
};

// graph.cpp:19-111
std::string make_graph(Layout const & layout, Node & root)
{
    root = Node{};
    root.max_bin_index = layout.top_level_max_bin_id;
    for (auto const & mb : layout.max_bins)
    {
        if (mb.previous_TB_indices.empty())
            return "layout: lower-level IBF without technical-bin path";
        Node * parent = &root;
        for (size_t k = 0; k + 1 < mb.previous_TB_indices.size(); ++k)
        {
            bool found = false;
            for (Node & child : parent->children)
                if (child.parent_bin_index == mb.previous_TB_indices[k])
                {
                    parent = &child;
                    found = true;
                    break;
                }
            if (!found)
                return "layout: lower-level IBF listed before its parent";
        }
        Node child{};
        child.parent_bin_index = mb.previous_TB_indices.back();
        child.max_bin_index = mb.id;
        parent->children.push_back(child);
        if (parent->max_bin_index == mb.previous_TB_indices.back())
            parent->favourite_child_idx = parent->children.size() - 1;
    }
    for (auto const & rec : layout.user_bins)
    {
        if (rec.number_of_technical_bins == 0)
            return "layout: user bin " + std::to_string(rec.idx) + " occupies no technical bin";
        if (rec.storage_TB_id >= kMaxLayoutBins || rec.number_of_technical_bins > kMaxLayoutBins - rec.storage_TB_id)
            return "layout: technical bin index beyond " + std::to_string(kMaxLayoutBins);
        Node * cur = &root;
        for (uint64_t bin : rec.previous_TB_indices)
        {
            if (bin >= kMaxLayoutBins)
                return "layout: technical bin index beyond " + std::to_string(kMaxLayoutBins);
            cur->number_of_technical_bins = std::max(cur->number_of_technical_bins, bin + 1);
            bool found = false;
            for (Node & child : cur->children)
                if (child.parent_bin_index == bin)
                {
                    cur = &child;
                    found = true;
                    break;
                }
            if (!found)
                return "layout: user bin refers to a merged bin without lower-level IBF";
        }
        cur->number_of_technical_bins =
            std::max(cur->number_of_technical_bins, rec.storage_TB_id + rec.number_of_technical_bins);
        if (rec.storage_TB_id == cur->max_bin_index)
            cur->remaining_records.insert(cur->remaining_records.begin(), rec);
        else
            cur->remaining_records.push_back(rec);
    }
    return {};
}

uint64_t divide_and_ceil(uint64_t a, uint64_t b) { return (a + (b - 1u)) / b; }

// The bounds a layout is checked against before its tables are sized from it.
bool node_bins_in_range(Node const & node, std::string & error)
{
    if (node.number_of_technical_bins == 0)
    {
        error = "layout: IBF without records";
        return false;
    }
    if (node.max_bin_index >= node.number_of_technical_bins)
    {
        error = "layout: fullest technical bin " + std::to_string(node.max_bin_index) + " beyond the IBF's " +
                std::to_string(node.number_of_technical_bins) + " technical bins";
        return false;
    }
    for (Node const & child : node.children)
        if (child.parent_bin_index >= node.number_of_technical_bins)
        {
            error = "layout: merged bin " + std::to_string(child.parent_bin_index) + " beyond the IBF's " +
                    std::to_string(node.number_of_technical_bins) + " technical bins";
            return false;
        }
    return true;
}

struct Builder
{
    Config const & c;
    BuildInput const & in;
    std::vector<double> corr;
    Plan & plan;
    std::string error;

    uint64_t union_size(std::vector<uint32_t> const & ubs) const
    {
        if (in.union_size)
            return in.union_size(ubs);
        uint64_t s = 0;
        for (uint32_t u : ubs)
            s += in.ub_size[u];
        return s;
    }

    // insert_into_ibf.cpp:28-47: chunks of ceil(n / bins) values, in order.
    void split_segments(uint32_t ibf, uint64_t first_tb, uint64_t number_of_bins, uint32_t ub)
    {
        uint64_t const n = in.ub_size[ub];
        if (n == 0)
            return;
        uint64_t const chunk = divide_and_ceil(n, number_of_bins);
        uint64_t tb = first_tb;
        for (uint64_t b = 0; b < n; b += chunk, ++tb)
            plan.segments.push_back({ibf, static_cast<uint32_t>(tb), ub, 0u, b, std::min(chunk, n - b)});
    }

    void whole_segments(uint32_t ibf, uint64_t tb, std::vector<uint32_t> const & ubs)
    {
        for (uint32_t u : ubs)
            if (in.ub_size[u])
                plan.segments.push_back({ibf, static_cast<uint32_t>(tb), u, 0u, 0u, in.ub_size[u]});
    }

    void fill_user_bins(PlannedIbf & p, Layout::UserBin const & rec)
    {
        for (uint64_t b = rec.storage_TB_id; b < rec.storage_TB_id + rec.number_of_technical_bins; ++b)
            p.user_bin[b] = rec.idx;
    }

    // hierarchical_interleaved_bloom_filter.cpp:40-171.  Returns the IBF index;
    // appends every user bin of the subtree to `subtree`.
    uint32_t build(Node const & node, bool is_root, std::vector<uint32_t> & subtree)
    {
        uint32_t const pos = static_cast<uint32_t>(plan.ibfs.size());
        plan.ibfs.emplace_back();
        if (!node_bins_in_range(node, error))
            return pos;
        uint64_t const ntb = node.number_of_technical_bins;
        {
            PlannedIbf & p = plan.ibfs[pos];
            p.next_ibf.assign(ntb, pos);
            p.user_bin.assign(ntb, kBinMerged);
        }
        // initialise_max_bin_kmers (:62-82)
        bool const max_merged = node.favourite_child_idx.has_value();
        uint64_t max_bin_tbs = 1;
        uint64_t max_elements = 0;
        std::vector<uint32_t> max_ubs;
        if (max_merged)
        {
            uint32_t const child = build(node.children[*node.favourite_child_idx], false, max_ubs);
            if (!error.empty())
                return pos;
            plan.ibfs[pos].next_ibf[node.max_bin_index] = child;
            max_elements = union_size(max_ubs);
        }
        else
        {
            if (node.remaining_records.empty())
            {
                error = "layout: IBF without records";
                return pos;
            }
            Layout::UserBin const & rec = node.remaining_records[0];
            if (rec.idx >= in.ub_size.size())
            {
                error = "layout: user bin index out of range";
                return pos;
            }
            max_bin_tbs = rec.number_of_technical_bins;
            max_elements = in.ub_size[rec.idx];
            fill_user_bins(plan.ibfs[pos], rec);
        }
        // construct_ibf (construct_ibf.cpp:26-64)
        if (max_bin_tbs >= corr.size())
        {
            error = "layout: user bin split over " + std::to_string(max_bin_tbs) + " technical bins, more than " +
                    "the top level's " + std::to_string(corr.size() - 1u);
            return pos;
        }
        uint64_t const per_bin = divide_and_ceil(max_elements, max_bin_tbs);
        double const fpr = max_merged ? c.relaxed_fpr : c.maximum_fpr;
        uint64_t const bits = bin_size_in_bits(fpr, c.number_of_hash_functions, per_bin);
        uint64_t const bin_size =
            max_bin_tbs == 1u ? bits : static_cast<uint64_t>(std::ceil(bits * corr[max_bin_tbs]));
        {
            PlannedIbf & p = plan.ibfs[pos];
            p.bins = ntb;
            p.bin_size = bin_size;
            p.hash_funs = c.number_of_hash_functions;
        }
        if (max_merged)
            whole_segments(pos, node.max_bin_index, max_ubs);
        else
            split_segments(pos, node.max_bin_index, max_bin_tbs, static_cast<uint32_t>(node.remaining_records[0].idx));
        subtree.insert(subtree.end(), max_ubs.begin(), max_ubs.end());

        // loop_over_children (:88-131), children in layout order
        for (size_t ci = 0; ci < node.children.size(); ++ci)
        {
            if (max_merged && ci == *node.favourite_child_idx)
                continue;
            std::vector<uint32_t> local;
            Node const & child = node.children[ci];
            uint32_t const cpos = build(child, false, local);
            if (!error.empty())
                return pos;
            plan.ibfs[pos].next_ibf[child.parent_bin_index] = cpos;
            whole_segments(pos, child.parent_bin_index, local);
            subtree.insert(subtree.end(), local.begin(), local.end());
        }
        // remaining records (:133-168)
        for (size_t i = max_merged ? 0u : 1u; i < node.remaining_records.size(); ++i)
        {
            Layout::UserBin const & rec = node.remaining_records[i];
            if (rec.idx >= in.ub_size.size())
            {
                error = "layout: user bin index out of range";
                return pos;
            }
            split_segments(pos, rec.storage_TB_id, rec.number_of_technical_bins, static_cast<uint32_t>(rec.idx));
            fill_user_bins(plan.ibfs[pos], rec);
            subtree.push_back(static_cast<uint32_t>(rec.idx));
        }
        if (!max_merged)
            subtree.push_back(static_cast<uint32_t>(node.remaining_records[0].idx));
        (void)is_root;
        return pos;
    }
};

} // namespace

std::string build_plan(Config const & c, Layout const & layout, BuildInput const & in, Plan & out)
{
    out = Plan{};
    if (in.ub_size.size() != c.number_of_user_bins)
        return "number of user-bin sizes differs from config::number_of_user_bins";
    for (uint64_t s : in.ub_size)
        if (s == 0)
            return "every user bin needs at least one value";
    Node root;
    if (std::string e = make_graph(layout, root); !e.empty())
        return e;
    Builder b{c, in, {}, out, {}};
    // build_index (:181-209): corrections for t_max = the root's technical bins.
    // Only the vector's length depends on t_max; cover every split count that can occur.
    b.corr = fpr_correction(c.maximum_fpr, c.number_of_hash_functions,
                            std::max<uint64_t>({root.number_of_technical_bins, c.tmax, 1u}));
    std::vector<uint64_t> seen(c.number_of_user_bins, 0);
    std::vector<uint32_t> all;
    b.build(root, true, all);
    if (!b.error.empty())
        return b.error;
    for (uint32_t u : all)
        ++seen[u];
    for (uint64_t u = 0; u < seen.size(); ++u)
        if (seen[u] != 1)
            return "layout does not place every user bin exactly once";
    for (auto const & p : out.ibfs)
        if (p.bin_size == 0)
            return "The size of a bin must be > 0.";
    for (Segment const & seg : out.segments)
        if (seg.ibf >= out.ibfs.size() || seg.tb >= out.ibfs[seg.ibf].bins)
            return "layout: a user bin's technical bins run past the end of its IBF";
    return {};
}

} // namespace hibf_amd::host
