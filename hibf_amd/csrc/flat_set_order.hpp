// The iteration order of the set the reference builds split user bins from.
//
// The reference fills a split user bin's technical bins with consecutive
// chunks of a robin_hood::unordered_flat_set<uint64_t> in its iteration order
// (src/build/insert_into_ibf.cpp:28-47; the set is robin_hood 3.11.5,
// include/hibf/contrib/robin_hood.hpp).  That order is the table's slot
// order, which follows from every insertion, growth and clear the set has
// seen.  FlatSetOrder keeps a table whose slots evolve exactly as that set's
// do (the same hash, probe rule, growth points, info-byte widths and reuse
// after clear) and reports its slot order.  It stores, beside each key, a tag
// (the key's input position) so a caller can map the order back to its data.
//
// Rules of the set, as robin_hood states them (Table in robin_hood.hpp):
// - hash(k) = murmur3 finaliser first half; then h *= multiplier, h ^= h >> 33;
//   home slot (h >> 5) & mask, info byte inc + ((h & 31) >> info_shift).
// - The table is sorted by (home, hash bits of the info byte descending); an
//   insertion walks while its info is below the slot's, passes equal infos
//   (comparing keys), then shifts the run up to the next free slot.
// - Slots: 2^k with a buffer of min(0.8 * 2^k, 255) behind; at most 80 % full.
// - An info byte that would pass 255 marks the table full; the next insertion
//   first halves the info increment (one hash bit fewer) while the increment
//   is above 2, else rehashes: with a new multiplier at the same size when
//   under half full, at double size otherwise.
// - clear() keeps the slots, the multiplier and the fill limit, and resets the
//   info increment.
#pragma once

#include <sys/mman.h>

#include <cstdint>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>

namespace hibf_amd::host
{

namespace detail
{
struct FreeDeleter
{
    void operator()(void * p) const noexcept { std::free(p); }
};

// Arrays of the set's table.  Large ones ask for 2 MiB pages: the table is
// probed at random addresses, and with 4 KiB pages nearly every probe of a
// large table also misses the TLB.
template <typename T>
std::unique_ptr<T[], FreeDeleter> table_array(uint64_t n, bool zero)
{
    constexpr uint64_t kHuge = 2ULL << 20;
    uint64_t const bytes = std::max<uint64_t>(n * sizeof(T), 1u);
    void * p = nullptr;
    if (bytes >= kHuge)
    {
        p = std::aligned_alloc(kHuge, (bytes + kHuge - 1u) / kHuge * kHuge);
        if (p)
            (void)madvise(p, bytes, MADV_HUGEPAGE);
    }
    else
    {
        p = std::malloc(bytes);
    }
    if (!p)
        throw std::bad_alloc();
    if (zero)
        std::memset(p, 0, bytes);
    return std::unique_ptr<T[], FreeDeleter>(static_cast<T *>(p));
}
} // namespace detail

class FlatSetOrder
{
public:
    FlatSetOrder() = default;
    FlatSetOrder(FlatSetOrder const &) = delete;
    FlatSetOrder & operator=(FlatSetOrder const &) = delete;
    FlatSetOrder(FlatSetOrder &&) = default;
    FlatSetOrder & operator=(FlatSetOrder &&) = default;

    uint64_t size() const { return n_; }
    uint64_t mask() const { return mask_; }

    // Inserts key unless present (the set's emplace); a new key keeps `tag`.
    // Returns whether the key was new.
    bool emplace(uint64_t const key, uint32_t const tag)
    {
        for (int attempt = 0; attempt < 256; ++attempt)
        {
            uint64_t idx;
            uint32_t info;
            key_to_idx(key, idx, info);
            if (mask_ != 0)
            {
                while (info < info_[idx])
                    next(info, idx);
                while (info == info_[idx])
                {
                    if (slots_[idx].key == key)
                        return false;
                    next(info, idx);
                }
            }
            if (n_ >= max_allowed_)
            {
                increase_size();
                continue;
            }
            uint64_t const insertion_idx = idx;
            uint32_t const insertion_info = info;
            if (insertion_info + info_inc_ > 0xFFu)
                max_allowed_ = 0;
            while (info_[idx] != 0)
                next(info, idx);
            if (idx != insertion_idx)
                shift_up(idx, insertion_idx);
            slots_[insertion_idx] = Slot{key, tag};
            info_[insertion_idx] = static_cast<uint8_t>(insertion_info);
            ++n_;
            return true;
        }
        throw std::overflow_error("robin_hood set emulation: insertion failed after 256 attempts");
    }

    // Issues a prefetch of the slots `key` will probe first.
    void prefetch(uint64_t const key) const
    {
        if (mask_ == 0)
            return;
        uint64_t idx;
        uint32_t info;
        key_to_idx(key, idx, info);
        __builtin_prefetch(info_.get() + idx, 1);
        __builtin_prefetch(slots_.get() + idx, 1);
    }

    void clear()
    {
        if (n_ == 0)
            return;
        n_ = 0;
        uint64_t const nb = with_buffer(mask_ + 1u);
        std::memset(info_.get(), 0, nb + 8u);
        info_[nb] = 1;
        info_inc_ = kInitialInfoInc;
        info_shift_ = 0;
    }

    // f(key, tag) for every element in the set's iteration order.
    template <typename F>
    void for_each(F && f) const
    {
        if (mask_ == 0)
            return;
        uint64_t const nb = with_buffer(mask_ + 1u);
        for (uint64_t i = 0; i < nb; ++i)
            if (info_[i] != 0)
                f(slots_[i].key, slots_[i].tag);
    }

private:
    struct Slot
    {
        uint64_t key;
        uint32_t tag;
    };

    static constexpr uint32_t kInitialInfoInc = 32;
    static constexpr uint64_t kInitialSlots = 8;

    static uint64_t max_allowed_for(uint64_t const slots)
    {
        return slots <= UINT64_MAX / 100u ? slots * 80u / 100u : (slots / 100u) * 80u;
    }
    static uint64_t with_buffer(uint64_t const slots)
    {
        uint64_t const m = max_allowed_for(slots);
        return slots + (m < 0xFFu ? m : 0xFFu);
    }

    void key_to_idx(uint64_t const key, uint64_t & idx, uint32_t & info) const
    {
        uint64_t h = key;
        h ^= h >> 33;
        h *= 0xFF51AFD7ED558CCDULL;
        h ^= h >> 33;
        h *= multiplier_;
        h ^= h >> 33;
        info = info_inc_ + static_cast<uint32_t>((h & 31u) >> info_shift_);
        idx = (h >> 5) & mask_;
    }

    void next(uint32_t & info, uint64_t & idx) const
    {
        ++idx;
        info += info_inc_;
    }

    void shift_up(uint64_t const start, uint64_t const insertion_idx)
    {
        for (uint64_t i = start; i != insertion_idx; --i)
        {
            slots_[i] = slots_[i - 1];
            info_[i] = static_cast<uint8_t>(info_[i - 1] + info_inc_);
            if (info_[i] + info_inc_ > 0xFFu)
                max_allowed_ = 0;
        }
    }

    void init_data(uint64_t const slots)
    {
        n_ = 0;
        mask_ = slots - 1u;
        max_allowed_ = max_allowed_for(slots);
        uint64_t const nb = with_buffer(slots);
        slots_ = detail::table_array<Slot>(nb, false);
        // nb info bytes, the sentinel and 7 bytes of padding, as the set's
        info_ = detail::table_array<uint8_t>(nb + 8u, true);
        info_[nb] = 1;
        info_inc_ = kInitialInfoInc;
        info_shift_ = 0;
    }

    bool try_increase_info()
    {
        if (info_inc_ <= 2u)
            return false;
        info_inc_ >>= 1;
        ++info_shift_;
        uint64_t const nb = with_buffer(mask_ + 1u);
        // the set halves its info bytes 8 at a time (the last word reaching into the padding)
        for (uint64_t i = 0; i < nb; i += 8u)
        {
            uint64_t v;
            std::memcpy(&v, info_.get() + i, 8u);
            v = (v >> 1) & 0x7F7F7F7F7F7F7F7FULL;
            std::memcpy(info_.get() + i, &v, 8u);
        }
        info_[nb] = 1;
        max_allowed_ = max_allowed_for(mask_ + 1u);
        return true;
    }

    void insert_move(uint64_t const key, uint32_t const tag)
    {
        if (max_allowed_ == 0 && !try_increase_info())
            throw std::overflow_error("robin_hood set emulation: overflow during rehash");
        uint64_t idx;
        uint32_t info;
        key_to_idx(key, idx, info);
        while (info <= info_[idx])
        {
            ++idx;
            info += info_inc_;
        }
        uint64_t const insertion_idx = idx;
        uint8_t const insertion_info = static_cast<uint8_t>(info);
        if (insertion_info + info_inc_ > 0xFFu)
            max_allowed_ = 0;
        while (info_[idx] != 0)
            next(info, idx);
        if (idx != insertion_idx)
            shift_up(idx, insertion_idx);
        slots_[insertion_idx] = Slot{key, tag};
        info_[insertion_idx] = insertion_info;
        ++n_;
    }

    void rehash(uint64_t const slots)
    {
        auto old = std::move(slots_);
        auto old_info = std::move(info_);
        uint64_t const old_nb = with_buffer(mask_ + 1u);
        init_data(slots);
        if (old && old_nb > 1u)
            for (uint64_t i = 0; i < old_nb; ++i)
                if (old_info[i] != 0)
                    insert_move(old[i].key, old[i].tag);
    }

    void increase_size()
    {
        if (mask_ == 0)
        {
            init_data(kInitialSlots);
            return;
        }
        uint64_t const max_allowed = max_allowed_for(mask_ + 1u);
        if (n_ < max_allowed && try_increase_info())
            return;
        if (n_ * 2u < max_allowed_for(mask_ + 1u))
        {
            multiplier_ += 0xC4CEB9FE1A85EC54ULL;
            rehash(mask_ + 1u);
        }
        else
        {
            rehash((mask_ + 1u) * 2u);
        }
    }

    std::unique_ptr<Slot[], detail::FreeDeleter> slots_;   // keys and tags
    std::unique_ptr<uint8_t[], detail::FreeDeleter> info_; // info bytes
    uint64_t mask_{0};
    uint64_t n_{0};
    uint64_t max_allowed_{0};
    uint64_t multiplier_{0xC4CEB9FE1A85EC53ULL};
    uint32_t info_inc_{kInitialInfoInc};
    uint32_t info_shift_{0};
};

} // namespace hibf_amd::host
