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


def expected_arena(dev, user_bins, keep_order: bool = False) -> np.ndarray:
    """Bit matrices the reference build (hierarchical_build + insert_into_ibf
    chunking) would set, given the tables the device build exported and the
    user bins' values in sorted-unique order (or in the given order when
    keep_order is set and the values are already distinct)."""
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
                chunk = -(-len(vals) // k)
                for c in range(0, len(vals), chunk):
                    ibf.emplace(vals[c:c + chunk], b + c // chunk)
                b = e
        w0 = int(t["word_offsets"][i])
        arena[w0:w0 + ibf.words.size] = ibf.words
    return arena
