// Serving loop for HIBF membership queries: the filter stays resident, requests
// arrive on the host at their own times, are admitted while their device
// memory fits a budget, gathered for a batching window, copied to the device
// and answered by the asynchronous membership launch on one of several
// streams, so batches overlap on the device.  Latency of a request is the time
// from its arrival to the moment its ids are back on the host.
//
// Query values in pinned host memory (hipHostMalloc, e.g. a receive ring) are
// copied to the device directly; pageable values go through pinned staging
// buffers, large batches with several host threads.
//
// The same loop serves reads: a request is then a DNA read whose bases are
// copied (one byte per base instead of eight per k-mer hash), hashed on the
// device and given the threshold uint16_t(fraction * number of k-mers) there.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <thread>
#include <vector>

#include "hibf_gpu.h"
#include "kernels.hpp"
#include "serving_common.hpp"

using namespace hibf_amd::serving;

namespace
{

// The loop.  Requests are elements of `eb` bytes (8: k-mer hashes with host thresholds; 1: read
// bases, hashed on the device as `rh` says (k-mers or minimisers), thresholds from its fraction).
int serve_impl(hibf_gpu_hibf hibf,
               hibf_gpu_serve_params const * params,
               char const * h_in,
               uint64_t const eb,
               uint64_t const * h_offsets,
               uint16_t const * h_thresholds,
               ReadHashing const & rh,
               double const * h_arrival_s,
               uint64_t nq,
               double * h_latency_s,
               uint64_t * h_ids,
               uint32_t * h_counts,
               hibf_gpu_serve_stats * stats)
{
    bool const reads = eb == 1u;
    if (!hibf || !params || !h_in || !h_offsets || (!reads && !h_thresholds) || !h_arrival_s || !h_latency_s || !h_ids
        || !h_counts)
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (reads && !(rh.fraction >= 0.0))
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "fraction must be >= 0");
    if (reads && !rh.mini && (rh.k == 0 || rh.k > 32))
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "k must be in [1, 32]");
    if (reads && rh.mini)
        if (int const rc = hibf_amd::minimiser_check(rh.shape, rh.window_size))
            return rc;
    hibf_gpu_serve_params const p = *params;
    if (p.streams == 0 || p.max_batch == 0 || p.results_per_query == 0 || p.max_batch_values == 0)
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT,
                            "streams, max_batch, max_batch_values and results_per_query must be > 0");
    if (p.streams > 64)
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "at most 64 streams");
    if (!(p.window_us >= 0.0) || p.window_us > 60e6)
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "window_us must be in [0, 6e7]");
    uint32_t const R = p.results_per_query;
    // device bytes of a request: its elements (+ its hashes for a read), offset, threshold, ids
    auto bytes_of = [&](uint64_t q)
    { return (h_offsets[q + 1] - h_offsets[q]) * (eb + (reads ? 8u : 0u)) + 34u + 8u * R; };
    for (uint64_t q = 0; q < nq; ++q)
    {
        if (h_offsets[q + 1] < h_offsets[q])
            return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "offsets decrease at request " + std::to_string(q));
        if (h_offsets[q + 1] - h_offsets[q] > p.max_batch_values)
            return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT,
                                "request " + std::to_string(q) + " is larger than max_batch_values");
        if (bytes_of(q) > p.budget_bytes)
            return serve_reject(HIBF_GPU_ERR_BUDGET, "request " + std::to_string(q) + " needs " +
                                                         std::to_string(bytes_of(q)) +
                                                         " bytes of device memory, more than the budget");
        if (!(h_arrival_s[q] >= 0.0) || (q && h_arrival_s[q] < h_arrival_s[q - 1]))
            return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT,
                                "arrival times must be >= 0 and non-decreasing (request " + std::to_string(q) + ")");
    }

    bool const pinned = host_pinned(h_in, h_offsets[nq] * eb);
    unsigned const copy_threads = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<Slot> slots(p.streams);
    int rc = HIBF_GPU_OK;
    auto cleanup = [&]()
    {
        for (Slot & s : slots)
            slot_free(s);
    };
#define SERVE_TRY(expr)                                                                                                \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t const e_ = (expr);                                                                                  \
        if (e_ != hipSuccess)                                                                                          \
        {                                                                                                              \
            (void)hipGetLastError();                                                                                   \
            cleanup();                                                                                                 \
            return serve_fail(e_, #expr);                                                                              \
        }                                                                                                              \
    } while (0)
    uint64_t const B = p.max_batch;
    uint64_t const V = p.max_batch_values;
    // Queues one batch (requests qb .. qb + nb - 1, nv elements) on a slot: offsets and thresholds staged,
    // the elements copied from the pinned input directly or through the slot's pinned staging buffer.
    auto enqueue = [&](Slot & s, uint64_t const qb, uint64_t const nb, uint64_t const nv) -> int
    {
        uint64_t const base = h_offsets[qb];
        for (uint64_t i = 0; i <= nb; ++i)
            s.h_offs[i] = h_offsets[qb + i] - base;
        if (!reads)
            std::memcpy(s.h_thr, h_thresholds + qb, nb * 2u);
        char const * src = h_in + base * eb;
        if (nv && !pinned)
        {
            copy_host(s.h_vals, src, nv * eb, copy_threads);
            src = reinterpret_cast<char const *>(s.h_vals);
        }
        return slot_launch(s, hibf, nb, nv, src, eb, R, reads, rh);
    };
    for (Slot & s : slots)
        if ((rc = slot_alloc(s, hibf, B, V, R, eb, reads, !pinned, reads && rh.mini)) != HIBF_GPU_OK)
        {
            cleanup();
            return rc;
        }

    // Warm the copy paths on every stream before the clock starts: the HIP runtime sets up a path the
    // first time it copies a size class, and that took about 10 ms inside the loop (a batch's latency).
    // Each slot copies every size from 64 bytes up to its buffers' capacity, in both directions.
    {
        for (Slot & s : slots)
        {
            void const * const h_src = pinned ? static_cast<void const *>(h_in) : static_cast<void const *>(s.h_vals);
            uint64_t const in_bytes = std::min<uint64_t>(V * eb, pinned ? h_offsets[nq] * eb : V * eb);
            uint64_t const out_bytes = B * R * 8u;
            for (uint64_t b = 64u;; b *= 4u)
            {
                uint64_t const bi = std::min(b, in_bytes), bo = std::min(b, out_bytes);
                if (bi)
                    SERVE_TRY(hipMemcpyAsync(s.d_vals, h_src, bi, hipMemcpyHostToDevice, s.stream));
                if (bo)
                    SERVE_TRY(hipMemcpyAsync(s.h_ids, s.d_ids, bo, hipMemcpyDeviceToHost, s.stream));
                if (b >= in_bytes && b >= out_bytes)
                    break;
            }
            SERVE_TRY(hipMemcpyAsync(s.d_offs, s.h_offs, (B + 1) * 8u, hipMemcpyHostToDevice, s.stream));
            SERVE_TRY(hipMemcpyAsync(s.d_thr, s.h_thr, B * 2u, hipMemcpyHostToDevice, s.stream));
            SERVE_TRY(hipMemcpyAsync(s.h_hs, s.d_hs, B * 8u, hipMemcpyDeviceToHost, s.stream));
            SERVE_TRY(hipStreamSynchronize(s.stream));
        }
        // A pinned input is read once by the device, chunk by chunk: the first device read of a region of
        // a freshly pinned buffer took about 10 ms in the loop (the region's first batch paid it).
        if (pinned && h_offsets[nq] > 0)
        {
            uint64_t const total = h_offsets[nq] * eb;
            uint64_t const chunk = V * eb;
            size_t c = 0;
            for (uint64_t off = 0; off < total; off += chunk, ++c)
            {
                Slot & s = slots[c % slots.size()];
                SERVE_TRY(hipMemcpyAsync(s.d_vals, h_in + off, std::min(chunk, total - off), hipMemcpyHostToDevice,
                                         s.stream));
            }
            for (Slot & s : slots)
                SERVE_TRY(hipStreamSynchronize(s.stream));
        }
        // Then many one-query batches (the first request, results discarded) queued on all streams at
        // once, as the loop queues them, so that the runtime's per-operation resources have grown to
        // the loop's depth before any request is timed.
        if (nq > 0 && h_offsets[1] - h_offsets[0] <= V)
        {
            // Every round writes the same staging contents, so rounds need not wait for each other.
            for (int round = 0; round < 64; ++round)
                for (Slot & s : slots)
                    if ((rc = enqueue(s, 0u, 1u, h_offsets[1] - h_offsets[0])) != HIBF_GPU_OK)
                    {
                        cleanup();
                        return rc;
                    }
        }
        for (Slot & s : slots)
        {
            SERVE_TRY(hipStreamSynchronize(s.stream));
            uint64_t v = 0, lk = 0, rb = 0;
            (void)hibf_gpu_agent_stats(s.agent, &v, &lk, &rb, 1);
        }
    }

    // Touch the result arrays before the clock starts: first-touch page faults
    // inside the loop would be charged to the requests that hit them.
    std::memset(h_latency_s, 0, nq * sizeof(double));
    std::memset(h_counts, 0, nq * sizeof(uint32_t));
    std::memset(h_ids, 0, nq * R * sizeof(uint64_t));

    uint64_t next = 0;          // next request to arrive
    uint64_t pending = 0;       // first admitted request not yet launched
    uint64_t completed = 0;
    uint64_t in_flight_bytes = 0;
    uint64_t max_in_flight = 0;
    uint64_t batches = 0;
    uint64_t overflowed = 0;
    double const window = p.window_us * 1e-6;
    auto const t0 = std::chrono::steady_clock::now();
    double last_done = 0.0;
    double longest_pass = 0.0;
    double longest_launch = 0.0;
    double pass_start = 0.0;
    while (completed < nq)
    {
        double const now = now_s(t0);
        longest_pass = std::max(longest_pass, now - pass_start);
        pass_start = now;
        // Completed batches: results back on the host.
        for (Slot & s : slots)
        {
            if (!s.busy)
                continue;
            hipError_t const e = hipEventQuery(s.done);
            if (e == hipErrorNotReady)
                continue;
            if (e != hipSuccess)
            {
                (void)hipGetLastError();
                cleanup();
                return serve_fail(e, "batch completion");
            }
            double const t = now_s(t0);
            for (uint64_t q = s.q0; q < s.q1; ++q)
            {
                uint64_t const i = q - s.q0;
                uint64_t const n = s.h_hc[i];
                uint64_t const keep = slot_kept_ids(s, i, s.q1 - s.q0, R);
                overflowed += (n > R || keep < std::min<uint64_t>(n, R));
                h_counts[q] = static_cast<uint32_t>(n);
                if (keep)
                    std::memcpy(h_ids + q * R, s.h_ids + s.h_hs[i], keep * 8u);
                h_latency_s[q] = t - h_arrival_s[q];
            }
            completed += s.q1 - s.q0;
            in_flight_bytes -= s.bytes;
            s.busy = false;
            last_done = t;
        }
        // Admission: arrived requests while their memory fits the budget.
        while (next < nq && h_arrival_s[next] <= now && in_flight_bytes + bytes_of(next) <= p.budget_bytes)
        {
            in_flight_bytes += bytes_of(next);
            ++next;
        }
        max_in_flight = std::max(max_in_flight, in_flight_bytes);
        // Launch: the oldest admitted request has waited a window, or a batch is full.
        if (pending < next)
        {
            uint64_t q1 = pending;
            uint64_t nv = 0;
            while (q1 < next && q1 - pending < B && nv + (h_offsets[q1 + 1] - h_offsets[q1]) <= V)
            {
                nv += h_offsets[q1 + 1] - h_offsets[q1];
                ++q1;
            }
            bool const full = q1 - pending >= B || q1 < next;
            bool const waited = now - h_arrival_s[pending] >= window;
            bool const drained = next == nq;
            if (full || waited || drained)
            {
                auto free_slot = std::find_if(slots.begin(), slots.end(), [](Slot const & s) { return !s.busy; });
                if (free_slot != slots.end())
                {
                    double const launch_start = now_s(t0);
                    Slot & s = *free_slot;
                    uint64_t const nb = q1 - pending;
                    if ((rc = enqueue(s, pending, nb, nv)) != HIBF_GPU_OK)
                    {
                        cleanup();
                        return rc;
                    }
                    s.q0 = pending;
                    s.q1 = q1;
                    s.bytes = 0;
                    for (uint64_t q = pending; q < q1; ++q)
                        s.bytes += bytes_of(q);
                    s.busy = true;
                    pending = q1;
                    ++batches;
                    longest_launch = std::max(longest_launch, now_s(t0) - launch_start);
                }
            }
        }
        if (next < nq && pending == next && std::all_of(slots.begin(), slots.end(), [](Slot const & s) { return !s.busy; }))
        {
            // Idle until the next arrival.
            double const wait = h_arrival_s[next] - now_s(t0);
            if (wait > 50e-6)
            {
                std::this_thread::sleep_for(std::chrono::duration<double>(wait - 20e-6));
                pass_start = now_s(t0); // an idle wait is not a stall
            }
        }
    }
    if (stats)
    {
        uint64_t lookups = 0, visits = 0, row_bytes = 0;
        for (Slot & s : slots)
        {
            uint64_t v = 0, lk = 0, rb = 0;
            (void)hibf_gpu_agent_stats(s.agent, &v, &lk, &rb, 0);
            lookups += lk;
            visits += v;
            row_bytes += rb;
        }
        stats->seconds = last_done - h_arrival_s[0];
        stats->batches = batches;
        stats->lookups = lookups;
        stats->row_bytes = row_bytes;
        stats->max_in_flight_bytes = max_in_flight;
        stats->overflowed = overflowed;
        stats->longest_pass_s = longest_pass;
        stats->longest_launch_s = longest_launch;
    }
    cleanup();
    return HIBF_GPU_OK;
#undef SERVE_TRY
}

} // namespace

extern "C" int hibf_gpu_hibf_serve(hibf_gpu_hibf hibf,
                                   hibf_gpu_serve_params const * params,
                                   uint64_t const * h_values,
                                   uint64_t const * h_offsets,
                                   uint16_t const * h_thresholds,
                                   double const * h_arrival_s,
                                   uint64_t nq,
                                   double * h_latency_s,
                                   uint64_t * h_ids,
                                   uint32_t * h_counts,
                                   hibf_gpu_serve_stats * stats)
{
    return serve_impl(hibf, params, reinterpret_cast<char const *>(h_values), 8u, h_offsets, h_thresholds,
                      ReadHashing{}, h_arrival_s, nq, h_latency_s, h_ids, h_counts, stats);
}

extern "C" int hibf_gpu_hibf_serve_reads(hibf_gpu_hibf hibf,
                                         hibf_gpu_serve_params const * params,
                                         uint8_t const * h_bases,
                                         uint64_t const * h_read_offsets,
                                         uint32_t k,
                                         double fraction,
                                         double const * h_arrival_s,
                                         uint64_t nq,
                                         double * h_latency_s,
                                         uint64_t * h_ids,
                                         uint32_t * h_counts,
                                         hibf_gpu_serve_stats * stats)
{
    ReadHashing rh;
    rh.k = k;
    rh.fraction = fraction;
    return serve_impl(hibf, params, reinterpret_cast<char const *>(h_bases), 1u, h_read_offsets, nullptr, rh,
                      h_arrival_s, nq, h_latency_s, h_ids, h_counts, stats);
}

extern "C" int hibf_gpu_hibf_serve_reads_minimisers(hibf_gpu_hibf hibf,
                                                    hibf_gpu_serve_params const * params,
                                                    uint8_t const * h_bases,
                                                    uint64_t const * h_read_offsets,
                                                    uint64_t shape,
                                                    uint32_t window_size,
                                                    uint64_t seed,
                                                    double fraction,
                                                    double const * h_arrival_s,
                                                    uint64_t nq,
                                                    double * h_latency_s,
                                                    uint64_t * h_ids,
                                                    uint32_t * h_counts,
                                                    hibf_gpu_serve_stats * stats)
{
    ReadHashing rh;
    rh.mini = true;
    rh.shape = shape;
    rh.window_size = window_size;
    rh.seed = seed;
    rh.fraction = fraction;
    return serve_impl(hibf, params, reinterpret_cast<char const *>(h_bases), 1u, h_read_offsets, nullptr, rh,
                      h_arrival_s, nq, h_latency_s, h_ids, h_counts, stats);
}
