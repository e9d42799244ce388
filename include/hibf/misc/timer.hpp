// seqan::hibf::serial_timer and seqan::hibf::concurrent_timer: the timers the
// reference's HIBF exposes for its build phases (misc/timer.hpp in the
// reference; members at hierarchical_interleaved_bloom_filter.hpp:283-290).
// Same member functions: start()/stop() add one steady_clock interval,
// operator+= adds another timer's total as one interval, in_seconds() is the
// total, max_in_seconds() the longest interval, avg_in_seconds() the total over
// the number of intervals.  The concurrent timer keeps its sums in atomics so
// that threads may add to it at once.
#pragma once

#include <atomic>
#include <cassert>
#include <chrono>
#include <cstdint>

namespace seqan::hibf
{

class concurrent_timer;

class serial_timer
{
public:
    using clock_t = std::chrono::steady_clock;
    using rep_t = clock_t::rep;

    void start() { started_ = clock_t::now(); }

    void stop() { add_ticks((clock_t::now() - started_).count()); }

    template <typename timer_t>
    void operator+=(timer_t const & other)
    {
        add_ticks(other.total_ticks());
    }

    double in_seconds() const { return seconds(total_); }
    double max_in_seconds() const { return seconds(longest_); }
    double avg_in_seconds() const
    {
        assert(intervals_ > 0u);
        return in_seconds() / static_cast<double>(intervals_);
    }

    rep_t total_ticks() const { return total_; }

    // Adds one interval of the given length (used for phases timed elsewhere).
    void add_seconds(double const s)
    {
        add_ticks(std::chrono::duration_cast<clock_t::duration>(std::chrono::duration<double>(s)).count());
    }

    constexpr bool operator==(serial_timer const &) const { return true; }

private:
    static double seconds(rep_t const t) { return std::chrono::duration<double>(clock_t::duration{t}).count(); }

    void add_ticks(rep_t const t)
    {
        total_ += t;
        longest_ = t > longest_ ? t : longest_;
        ++intervals_;
    }

    clock_t::time_point started_{};
    rep_t total_{};
    rep_t longest_{};
    uint64_t intervals_{};
};

class concurrent_timer
{
public:
    using clock_t = std::chrono::steady_clock;
    using rep_t = clock_t::rep;

    concurrent_timer() = default;
    concurrent_timer(concurrent_timer const & o) { copy_from(o); }
    concurrent_timer & operator=(concurrent_timer const & o)
    {
        copy_from(o);
        return *this;
    }
    concurrent_timer(concurrent_timer && o) noexcept { copy_from(o); }
    concurrent_timer & operator=(concurrent_timer && o) noexcept
    {
        copy_from(o);
        return *this;
    }

    void start() { started_ = clock_t::now(); }

    void stop() { add_ticks((clock_t::now() - started_).count()); }

    template <typename timer_t>
    void operator+=(timer_t const & other)
    {
        add_ticks(other.total_ticks());
    }

    double in_seconds() const { return seconds(total_.load()); }
    double max_in_seconds() const { return seconds(longest_.load()); }
    double avg_in_seconds() const
    {
        assert(intervals_.load() > 0u);
        return in_seconds() / static_cast<double>(intervals_.load());
    }

    rep_t total_ticks() const { return total_.load(); }

    // Adds one interval of the given length (used for phases timed elsewhere).
    void add_seconds(double const s)
    {
        add_ticks(std::chrono::duration_cast<clock_t::duration>(std::chrono::duration<double>(s)).count());
    }

    constexpr bool operator==(concurrent_timer const &) const { return true; }

private:
    static double seconds(rep_t const t) { return std::chrono::duration<double>(clock_t::duration{t}).count(); }

    void add_ticks(rep_t const t)
    {
        total_ += t;
        for (rep_t seen = longest_.load(); seen < t && !longest_.compare_exchange_weak(seen, t);)
        {
        }
        ++intervals_;
    }

    void copy_from(concurrent_timer const & o)
    {
        started_ = o.started_;
        total_ = o.total_.load();
        longest_ = o.longest_.load();
        intervals_ = o.intervals_.load();
    }

    clock_t::time_point started_{};
    alignas(64) std::atomic<rep_t> total_{};
    alignas(64) std::atomic<rep_t> longest_{};
    alignas(64) std::atomic<uint64_t> intervals_{};
};

} // namespace seqan::hibf
