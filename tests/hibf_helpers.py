"""Test helpers: turn a device HIBF into the oracle's form, and recompute the
bit matrices its build must produce from the exported routing tables."""
from __future__ import annotations

import numpy as np

import oracle as O


def oracle_of(dev) -> O.Hibf:
    t = dev.tables()
    geoms = [O.geom(b, s, h) for (b, s, h) in t["geoms"]]
    arena = dev.download_arena()
    return O.Hibf(geoms, arena, t["word_offsets"], t["tb_offsets"], t["next_ibf"], t["user_bin"], t["n_user_bins"])


def subtree_user_bins(t, ibf: int) -> list[int]:
    """User bins below IBF `ibf` (regular bins of it and of every merged descendant)."""
    out, stack = [], [ibf]
    while stack:
        i = stack.pop()
        lo, hi = int(t["tb_offsets"][i]), int(t["tb_offsets"][i + 1])
        for b in range(lo, hi):
            u = int(t["user_bin"][b])
            if u == O.BIN_MERGED:
                stack.append(int(t["next_ibf"][b]))
            elif u != O.BIN_DELETED:
                out.append(u)
    return sorted(set(out))


class _Node:
    def __init__(self, parent_bin_index=0, max_bin_index=0):
        self.children = []
        self.parent_bin_index = parent_bin_index
        self.max_bin_index = max_bin_index
        self.ntb = 0
        self.fav = None
        self.records = []  # (user bin, storage technical bin, number of technical bins)


def layout_graph(text: str) -> _Node:
    """The reference's layout graph (src/layout/graph.cpp:19-125) of a layout
    in its text format (src/layout/layout.cpp:125-136): header lines name each
    lower-level IBF by its technical-bin path and fullest bin; each user bin
    line gives its path, storage bin and number of technical bins.  Records
    keep file order, the one stored in the IBF's fullest bin moved first."""
    root = _Node()
    for line in text.splitlines():
        if line.startswith("#TOP_LEVEL_IBF"):
            root.max_bin_index = int(line.split(":")[1])
        elif line.startswith("#LOWER_LEVEL_IBF_"):
            path = [int(x) for x in line.split(" ")[0][len("#LOWER_LEVEL_IBF_"):].split(";")]
            parent = root
            for b in path[:-1]:
                parent = next(c for c in parent.children if c.parent_bin_index == b)
            child = _Node(path[-1], int(line.split(":")[1]))
            parent.children.append(child)
            if parent.max_bin_index == path[-1]:
                parent.fav = len(parent.children) - 1
        elif line and not line.startswith("#") and not line.startswith("@"):
            ub, bins, counts = line.split("\t")
            bins = [int(x) for x in bins.split(";")]
            counts = [int(x) for x in counts.split(";")]
            node = root
            for b in bins[:-1]:
                node.ntb = max(node.ntb, b + 1)
                node = next(c for c in node.children if c.parent_bin_index == b)
            node.ntb = max(node.ntb, bins[-1] + counts[-1])
            rec = (int(ub), bins[-1], counts[-1])
            if bins[-1] == node.max_bin_index:
                node.records.insert(0, rec)
            else:
                node.records.append(rec)
    return root


def reference_split_chunks(layout_text: str, insert, device_index) -> dict:
    """Chunk number of every value of every split user bin, as the reference's
    build assigns them: the whole of hierarchical_build
    (src/hierarchical_interleaved_bloom_filter.cpp:40-171, construct_ibf.cpp:
    59-61, insert_into_ibf.cpp:28-47, update_parent_kmers.hpp) replayed on the
    oracle's restatement of its k-mer set (robin_hood::unordered_flat_set).
    insert(set, u) adds user bin u's values in input order; device_index(u,
    keys) gives the keys' positions in the device copy of user bin u.
    Returns {u: chunk numbers indexed by device position}."""
    out = {}

    def chunks(kmers, u, bins):
        order = kmers.order()
        size = -(-order.size // bins)
        c = np.empty(order.size, dtype=np.int64)
        c[np.asarray(device_index(u, order), dtype=np.int64)] = np.arange(order.size) // size
        out[u] = c

    def build(node, is_root, parent):
        kmers = O.FlatSet()
        if node.fav is not None:  # initialise_max_bin_kmers: the favourite child fills kmers
            build(node.children[node.fav], False, kmers)
        else:
            u, _, bins = node.records[0]
            insert(kmers, u)
            if bins > 1:
                chunks(kmers, u, bins)
        if not is_root:  # construct_ibf
            parent.insert_set(kmers)
        kmers.clear()
        for ci, child in enumerate(node.children):  # loop_over_children
            if ci == node.fav:
                continue
            local = O.FlatSet()
            build(child, False, local)
            if not is_root:
                parent.insert_set(local)
        for u, _, bins in node.records[0 if node.fav is not None else 1:]:
            if is_root and bins == 1:  # inserted into the IBF directly
                continue
            insert(kmers, u)
            if bins > 1:
                chunks(kmers, u, bins)
            if not is_root:
                parent.insert_set(kmers)
            kmers.clear()

    build(layout_graph(layout_text), True, None)
    return out


def node_split_chunks(layout_text: str, path, insert, device_index, max_splits: int | None = None) -> dict:
    """reference_split_chunks for the split user bins of one IBF only, the one
    at technical-bin path `path` (() = the root): its k-mer set's history is
    its fullest bin's content (the favourite child's whole subtree, replayed
    in full, when that bin is merged) and then its remaining user bins in
    order; its other children use sets of their own and cannot change it.
    With max_splits, stops after that many split user bins."""
    out = {}
    root = layout_graph(layout_text)
    node = root
    for b in path:
        node = next(c for c in node.children if c.parent_bin_index == b)
    is_root = len(path) == 0

    def chunks(kmers, u, bins):
        order = kmers.order()
        size = -(-order.size // bins)
        c = np.empty(order.size, dtype=np.int64)
        c[np.asarray(device_index(u, order), dtype=np.int64)] = np.arange(order.size) // size
        out[u] = c

    def subtree_into(n, parent):  # hierarchical_build of a non-root node, its content into parent
        kmers = O.FlatSet()
        if n.fav is not None:
            subtree_into(n.children[n.fav], kmers)
        else:
            insert(kmers, n.records[0][0])
        parent.insert_set(kmers)
        kmers.clear()
        for ci, child in enumerate(n.children):
            if ci == n.fav:
                continue
            local = O.FlatSet()
            subtree_into(child, local)
            parent.insert_set(local)
        for u, _, _ in n.records[0 if n.fav is not None else 1:]:
            insert(kmers, u)
            parent.insert_set(kmers)
            kmers.clear()

    kmers = O.FlatSet()
    if node.fav is not None:
        subtree_into(node.children[node.fav], kmers)
    else:
        u, _, bins = node.records[0]
        insert(kmers, u)
        if bins > 1:
            chunks(kmers, u, bins)
    kmers.clear()
    last = max([i for i, r in enumerate(node.records) if r[2] > 1], default=-1)
    for i, (u, _, bins) in enumerate(node.records):
        if i == 0 and node.fav is None:
            continue
        if i > last:
            break
        if is_root and bins == 1:
            continue
        if max_splits is not None and len(out) >= max_splits:
            break
        insert(kmers, u)
        if bins > 1:
            chunks(kmers, u, bins)
        kmers.clear()
    return out


def synthetic_split_chunks(dev, seed: int, sizes) -> dict:
    """reference_split_chunks of an HIBF built from synthetic user bins
    (value i of user bin u = synth_value(seed, u, i) for i < sizes[u],
    inserted in order of i; the device position of value i is i)."""
    from hibf_amd.synth import synth_values

    def insert(s, u):
        s.insert_synth(seed, u, 0, int(sizes[u]))

    def device_index(u, keys):
        vals = synth_values(seed, u, np.arange(int(sizes[u]), dtype=np.uint64))
        # keys are the set's iteration order of exactly these values: match the two sorted orders
        assert keys.size == vals.size
        out = np.empty(keys.size, dtype=np.int64)
        out[np.argsort(keys)] = np.argsort(vals)
        return out

    return reference_split_chunks(dev.layout_text(), insert, device_index)


def explicit_split_chunks(dev, user_bins) -> dict:
    """reference_split_chunks of an HIBF built from explicit user bins (values
    in their given order, duplicates included; the device holds each user
    bin's sorted distinct values)."""
    sets = [np.unique(np.asarray(u, dtype=np.uint64)) for u in user_bins]

    def insert(s, u):
        s.insert_values(np.asarray(user_bins[u], dtype=np.uint64))

    def device_index(u, keys):
        return np.searchsorted(sets[u], keys)

    return reference_split_chunks(dev.layout_text(), insert, device_index)


def expected_arena(dev, user_bins, keep_order: bool = False, split_chunks: dict | None = None) -> np.ndarray:
    """Bit matrices the reference build (hierarchical_build + insert_into_ibf
    chunking) would set, given the tables the device build exported and the
    user bins' values: in sorted distinct order, or in the given order when
    keep_order is set and the values are already distinct.  A split user bin
    is chunked by split_chunks (chunk number per value, as
    reference_split_chunks gives them) when given, else by ranges of that
    order."""
    t = dev.tables()
    sets = [np.asarray(u, dtype=np.uint64) if keep_order else np.unique(np.asarray(u, dtype=np.uint64))
            for u in user_bins]
    arena = np.zeros(int(t["n_words"]), dtype=np.uint64)
    for i, (bins, bin_size, h) in enumerate(t["geoms"]):
        ibf = O.Ibf(bins, bin_size, h)
        lo = int(t["tb_offsets"][i])
        ub = [int(x) for x in t["user_bin"][lo:lo + bins]]
        nx = [int(x) for x in t["next_ibf"][lo:lo + bins]]
        b = 0
        while b < bins:
            u = ub[b]
            if u == O.BIN_MERGED:
                for v in subtree_user_bins(t, nx[b]):
                    ibf.emplace(sets[v], b)
                b += 1
            elif u == O.BIN_DELETED:
                b += 1
            else:
                e = b
                while e < bins and ub[e] == u:
                    e += 1
                k = e - b
                vals = sets[u]
                if k > 1 and split_chunks is not None:
                    c = split_chunks[u]
                    assert c.size == vals.size and c.max() < k
                    ibf.emplace(vals, b + c.astype(np.uint64))
                else:
                    chunk = -(-len(vals) // k)
                    for c0 in range(0, len(vals), chunk):
                        ibf.emplace(vals[c0:c0 + chunk], b + c0 // chunk)
                b = e
        w0 = int(t["word_offsets"][i])
        arena[w0:w0 + ibf.words.size] = ibf.words
    return arena
