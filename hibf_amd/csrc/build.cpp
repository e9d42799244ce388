// HIBF build plan from a layout: restates the reference's layout graph
// (src/layout/graph.cpp:19-125), construct_ibf (src/build/construct_ibf.cpp:26-64),
// insert_into_ibf's split-bin chunking (src/build/insert_into_ibf.cpp:28-69) and
// hierarchical_build (src/hierarchical_interleaved_bloom_filter.cpp:40-179).
// Instead of inserting values, it records which value ranges go to which
// technical bin; the device bulk emplace then sets the bits.
#include "build.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <mutex>
#include <optional>
#include <string>
#include <thread>

#include "flat_set_order.hpp"

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
    std::vector<int64_t> chunked_segment_of{}; // per user bin: its chunked segment, -1 if none
    uint64_t max_split{0};

    uint64_t union_size(std::vector<uint32_t> const & ubs) const
    {
        if (in.union_size)
            return in.union_size(ubs);
        uint64_t s = 0;
        for (uint32_t u : ubs)
            s += in.ub_size[u];
        return s;
    }

    // insert_into_ibf.cpp:28-47: chunks of ceil(n / bins) values.  With a
    // value source the chunks follow the reference's set order (one chunked
    // segment, filled by SplitOrder); otherwise they are ranges of the
    // device order.
    void split_segments(uint32_t ibf, uint64_t first_tb, uint64_t number_of_bins, uint32_t ub)
    {
        uint64_t const n = in.ub_size[ub];
        if (n == 0)
            return;
        if (in.values && number_of_bins > 1u)
        {
            chunked_segment_of[ub] = static_cast<int64_t>(plan.segments.size());
            plan.segments.push_back(
                {ibf, static_cast<uint32_t>(first_tb), ub, kChunked | static_cast<uint32_t>(number_of_bins << 8), 0u, n});
            max_split = std::max(max_split, number_of_bins);
            return;
        }
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

// Replays the reference's use of its k-mer sets over the layout graph
// (hierarchical_build, src/hierarchical_interleaved_bloom_filter.cpp:40-171;
// construct_ibf, src/build/construct_ibf.cpp:59-61; insert_into_ibf and
// update_parent_kmers) far enough to know the order in which every split user
// bin is chunked, and writes each split bin's chunk numbers.
//
// Per IBF the reference keeps one set, `kmers`: it receives the fullest bin's
// content (a user bin, or the favourite child's whole subtree through
// update_parent_kmers), is cleared, and then receives each remaining user bin
// that goes through a set (at the root only split ones) in layout order,
// cleared after each.  clear() keeps the set's table, so a user bin's order
// depends on what the set held before.  A set whose content the parent also
// receives (update_parent_kmers, every IBF but the root) passes it on in its
// iteration order.  Only what can change some split bin's order is replayed:
// an IBF's records up to its last split one, and an IBF's whole subtree when
// its parent's set order depends on it.
class SplitOrder
{
public:
    SplitOrder(BuildInput const & in, Plan & plan, std::vector<int64_t> const & chunked_segment_of) :
        in_{in},
        plan_{plan},
        seg_of_{chunked_segment_of}
    {}

    std::string run(Node const & root)
    {
        collect(root, true, false);
        // largest tasks first, so that the longest replays start early
        std::vector<std::pair<uint64_t, size_t>> order;
        for (size_t t = 0; t < tasks_.size(); ++t)
            order.push_back({tasks_[t].work, t});
        std::sort(order.begin(), order.end(), std::greater<>());
        std::atomic<size_t> next{0};
        unsigned const threads = std::max(1u, std::min<unsigned>(in_.threads, static_cast<unsigned>(order.size())));
        auto worker = [&]()
        {
            for (size_t k; (k = next.fetch_add(1)) < order.size();)
            {
                Task const & t = tasks_[order[k].second];
                try
                {
                    replay(*t.node, t.is_root, nullptr, false);
                }
                catch (std::exception const & e)
                {
                    std::lock_guard<std::mutex> lock(error_mutex_);
                    if (error_.empty())
                        error_ = std::string("split-bin order: ") + e.what();
                }
            }
        };
        std::vector<std::thread> pool;
        for (unsigned i = 1; i < threads; ++i)
            pool.emplace_back(worker);
        worker();
        for (auto & th : pool)
            th.join();
        return error_;
    }

private:
    struct Task
    {
        Node const * node;
        bool is_root;
        uint64_t work; // values the replay inserts, for scheduling
    };

    static uint64_t start_of(Node const & n) { return n.favourite_child_idx ? 0u : 1u; }

    // Index of the last remaining record, from start_of on, that is split (-1: none).
    static int64_t last_split(Node const & n)
    {
        int64_t last = -1;
        for (size_t i = start_of(n); i < n.remaining_records.size(); ++i)
            if (n.remaining_records[i].number_of_technical_bins > 1u)
                last = static_cast<int64_t>(i);
        return last;
    }

    uint64_t subtree_values(Node const & n) const
    {
        uint64_t v = 0;
        for (auto const & r : n.remaining_records)
            v += in_.ub_size[r.idx];
        for (Node const & c : n.children)
            v += subtree_values(c);
        return v;
    }

    // need_out: the node's content goes, in its sets' iteration orders, into a
    // set of its parent whose order matters; such a node is replayed inside its
    // parent's task.  Every other node is a task of its own.
    void collect(Node const & n, bool is_root, bool need_out)
    {
        bool const need_after = last_split(n) >= 0 || need_out;
        if (!need_out)
        {
            uint64_t work = 0;
            int64_t const last = last_split(n);
            if (!n.favourite_child_idx && !n.remaining_records.empty() &&
                (n.remaining_records[0].number_of_technical_bins > 1u || need_after))
                work += in_.ub_size[n.remaining_records[0].idx];
            if (n.favourite_child_idx && need_after)
                work += 2u * subtree_values(n.children[*n.favourite_child_idx]);
            for (int64_t i = static_cast<int64_t>(start_of(n)); i <= last; ++i)
                work += in_.ub_size[n.remaining_records[i].idx];
            tasks_.push_back({&n, is_root, work});
        }
        for (size_t ci = 0; ci < n.children.size(); ++ci)
        {
            bool const fav = n.favourite_child_idx && ci == *n.favourite_child_idx;
            collect(n.children[ci], false, fav ? need_after : need_out);
        }
    }

    void replay(Node const & n, bool is_root, FlatSetOrder * parent, bool need_out)
    {
        FlatSetOrder kmers;
        int64_t const last = last_split(n);
        bool const need_after = last >= 0 || need_out;
        // initialise_max_bin_kmers
        if (n.favourite_child_idx)
        {
            if (need_after)
                replay(n.children[*n.favourite_child_idx], false, &kmers, true);
        }
        else
        {
            Layout::UserBin const & r0 = n.remaining_records[0];
            if (r0.number_of_technical_bins > 1u || need_after)
            {
                insert_user_bin(kmers, static_cast<uint32_t>(r0.idx));
                if (r0.number_of_technical_bins > 1u)
                    assign_chunks(kmers, static_cast<uint32_t>(r0.idx), r0.number_of_technical_bins);
            }
        }
        if (need_out)
            insert_set(*parent, kmers); // construct_ibf's update_parent_kmers
        kmers.clear();
        // loop_over_children: the other children in index order, each into a fresh set
        if (need_out)
            for (size_t ci = 0; ci < n.children.size(); ++ci)
            {
                if (n.favourite_child_idx && ci == *n.favourite_child_idx)
                    continue;
                FlatSetOrder local;
                replay(n.children[ci], false, &local, true);
                insert_set(*parent, local);
            }
        // the remaining records
        for (size_t i = start_of(n); i < n.remaining_records.size(); ++i)
        {
            Layout::UserBin const & r = n.remaining_records[i];
            if (is_root && r.number_of_technical_bins == 1u)
                continue; // inserted into the IBF directly, no set
            if (!need_out && static_cast<int64_t>(i) > last)
                break;
            insert_user_bin(kmers, static_cast<uint32_t>(r.idx));
            if (r.number_of_technical_bins > 1u)
                assign_chunks(kmers, static_cast<uint32_t>(r.idx), r.number_of_technical_bins);
            if (need_out)
                insert_set(*parent, kmers);
            kmers.clear();
        }
    }

    // compute_kmers: input_fn's values of user bin u into the set.
    void insert_user_bin(FlatSetOrder & s, uint32_t u) const
    {
        constexpr uint64_t kBuf = 4096;
        constexpr uint64_t kAhead = 16;
        std::vector<uint64_t> vals(kBuf);
        std::vector<uint32_t> tags(kBuf);
        for (uint64_t pos = 0;;)
        {
            uint64_t const got = in_.values->read(u, pos, kBuf, vals.data(), tags.data());
            if (got == 0)
                break;
            for (uint64_t k = 0; k < got; ++k)
            {
                if (k + kAhead < got)
                    s.prefetch(vals[k + kAhead]);
                s.emplace(vals[k], tags[k]);
            }
            pos += got;
        }
    }

    // update_parent_kmers: dst.insert(src.begin(), src.end()).
    static void insert_set(FlatSetOrder & dst, FlatSetOrder const & src)
    {
        src.for_each([&](uint64_t key, uint32_t) { dst.emplace(key, 0u); });
    }

    // insert_into_ibf: chunk c of the set's iteration order goes to technical bin first + c.
    void assign_chunks(FlatSetOrder const & s, uint32_t u, uint64_t bins)
    {
        if (s.size() != in_.ub_size[u])
            throw std::runtime_error("user bin " + std::to_string(u) + " holds " + std::to_string(s.size()) +
                                     " distinct values, not the " + std::to_string(in_.ub_size[u]) + " planned");
        int64_t const seg = seg_of_[u];
        if (seg < 0)
            throw std::runtime_error("user bin " + std::to_string(u) + " has no chunked segment");
        uint64_t const chunk = divide_and_ceil(s.size(), bins);
        uint64_t const base = plan_.segments[static_cast<size_t>(seg)].begin;
        uint8_t * const ids = plan_.chunk_ids.data();
        uint32_t const w = plan_.chunk_id_bytes;
        uint64_t rank = 0;
        s.for_each(
            [&](uint64_t, uint32_t tag)
            {
                uint64_t const c = rank++ / chunk;
                if (w == 1u)
                    ids[base + tag] = static_cast<uint8_t>(c);
                else
                {
                    ids[2u * (base + tag)] = static_cast<uint8_t>(c);
                    ids[2u * (base + tag) + 1u] = static_cast<uint8_t>(c >> 8);
                }
            });
    }

    BuildInput const & in_;
    Plan & plan_;
    std::vector<int64_t> const & seg_of_;
    std::vector<Task> tasks_;
    std::mutex error_mutex_;
    std::string error_;
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
    b.chunked_segment_of.assign(c.number_of_user_bins, -1);
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
    if (b.max_split > 65536u)
        return "split user bins over more than 65,536 technical bins are not supported";
    for (Segment const & seg : out.segments)
    {
        uint64_t const last_tb = seg.flags & kChunked ? seg.tb + (seg.flags >> 8) - 1u : seg.tb;
        if (seg.ibf >= out.ibfs.size() || last_tb >= out.ibfs[seg.ibf].bins)
            return "layout: a user bin's technical bins run past the end of its IBF";
    }
    if (in.values && b.max_split > 0)
    {
        out.chunk_id_bytes = b.max_split <= 256u ? 1u : 2u;
        uint64_t total = 0;
        for (Segment & seg : out.segments)
            if (seg.flags & kChunked)
            {
                seg.begin = total;
                total += seg.count;
            }
        out.chunk_ids.assign(total * out.chunk_id_bytes, 0);
        SplitOrder order{in, out, b.chunked_segment_of};
        if (std::string e = order.run(root); !e.empty())
            return e;
    }
    return {};
}

} // namespace hibf_amd::host
