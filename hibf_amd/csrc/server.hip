// Online serving of HIBF membership queries: callers submit requests while
// the loop runs (from any number of host threads) and collect the results as
// batches finish.  The loop is the one of the trace-replay form
// (serving.hip): requests are admitted in submission order while their device
// memory fits the budget, gathered until the oldest has waited the batching
// window or a batch is full, staged into a stream's pinned buffers and
// answered by the asynchronous membership launch, several streams in flight.
// Latency is the time from submission to the result being ready to poll.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hibf_gpu.h"
#include "kernels.hpp"
#include "serving_common.hpp"

using namespace hibf_amd::serving;

namespace
{

using clock_type = std::chrono::steady_clock;

struct Request
{
    uint64_t ticket{};
    std::vector<uint64_t> values;
    uint16_t threshold{};
    double submitted_s{};
    uint64_t bytes{}; // device memory while admitted
};

struct Batched
{
    uint64_t ticket;
    double submitted_s;
    uint64_t bytes;
};

struct Finished
{
    hibf_gpu_server_result result;
    std::vector<uint64_t> ids;
};

} // namespace

struct hibf_gpu_server_s
{
    hibf_gpu_hibf hibf{};
    hibf_gpu_serve_params p{};
    int device{0};
    clock_type::time_point t0{clock_type::now()};
    std::vector<Slot> slots;
    std::vector<std::vector<Batched>> batch_of; // per slot: the requests of its batch
    std::thread loop;

    std::mutex m;
    std::condition_variable cv_in;  // the loop waits for submissions
    std::condition_variable cv_out; // pollers wait for results
    std::deque<Request> incoming;   // submitted, not yet admitted
    std::deque<Finished> finished;  // answered, not yet polled
    bool stopping{false};
    int error{HIBF_GPU_OK};
    std::string error_text;
    uint64_t next_ticket{0};
    uint64_t outstanding{0}; // submitted and not yet answered

    // statistics (under m)
    uint64_t batches{0};
    uint64_t overflowed{0};
    uint64_t max_in_flight{0};
    double first_submit_s{-1.0};
    double last_done_s{0.0};
    double longest_pass_s{0.0};
    double longest_launch_s{0.0};

    double now() const { return std::chrono::duration<double>(clock_type::now() - t0).count(); }
    uint64_t bytes_of(uint64_t n) const { return n * 8u + 34u + 8u * p.results_per_query; }

    void fail(int status)
    {
        std::lock_guard<std::mutex> lock(m);
        if (error == HIBF_GPU_OK)
        {
            error = status;
            error_text = hibf_gpu_last_error();
        }
        cv_out.notify_all();
    }

    void run();
};

void hibf_gpu_server_s::run()
{
    if (hipSetDevice(device) != hipSuccess)
    {
        (void)hipGetLastError();
        fail(serve_fail(hipErrorInvalidDevice, "set device"));
        return;
    }
    uint64_t const B = p.max_batch, V = p.max_batch_values;
    uint32_t const R = p.results_per_query;
    double const window = p.window_us * 1e-6;
    std::deque<Request> admitted;
    uint64_t in_flight = 0;
    double pass_start = now();
    for (;;)
    {
        double const t_pass = now();
        {
            std::lock_guard<std::mutex> lock(m);
            longest_pass_s = std::max(longest_pass_s, t_pass - pass_start);
        }
        pass_start = t_pass;
        // Finished batches: results out to the pollers.
        for (size_t si = 0; si < slots.size(); ++si)
        {
            Slot & s = slots[si];
            if (!s.busy)
                continue;
            hipError_t const e = hipEventQuery(s.done);
            if (e == hipErrorNotReady)
                continue;
            if (e != hipSuccess)
            {
                (void)hipGetLastError();
                fail(serve_fail(e, "batch completion"));
                return;
            }
            std::vector<Batched> & reqs = batch_of[si];
            uint64_t const nb = reqs.size();
            double const t = now();
            std::vector<Finished> done(nb);
            uint64_t over = 0;
            for (uint64_t i = 0; i < nb; ++i)
            {
                uint64_t const n = s.h_hc[i];
                uint64_t const keep = slot_kept_ids(s, i, nb, R);
                over += (n > R || keep < std::min<uint64_t>(n, R));
                done[i].result = hibf_gpu_server_result{reqs[i].ticket, static_cast<uint32_t>(n),
                                                        static_cast<uint32_t>(keep), t - reqs[i].submitted_s};
                done[i].ids.assign(s.h_ids + s.h_hs[i], s.h_ids + s.h_hs[i] + keep);
                in_flight -= reqs[i].bytes;
            }
            reqs.clear();
            s.busy = false;
            {
                std::lock_guard<std::mutex> lock(m);
                for (Finished & f : done)
                    finished.push_back(std::move(f));
                outstanding -= nb;
                overflowed += over;
                last_done_s = t;
            }
            cv_out.notify_all();
        }
        // Admission in submission order while the memory fits the budget.
        bool stop;
        {
            std::lock_guard<std::mutex> lock(m);
            while (!incoming.empty() && in_flight + incoming.front().bytes <= p.budget_bytes)
            {
                in_flight += incoming.front().bytes;
                admitted.push_back(std::move(incoming.front()));
                incoming.pop_front();
            }
            max_in_flight = std::max(max_in_flight, in_flight);
            stop = stopping && incoming.empty();
        }
        // Launch: the oldest admitted request has waited a window, a batch is full, or the loop is stopping.
        if (!admitted.empty())
        {
            uint64_t nb = 0, nv = 0;
            while (nb < admitted.size() && nb < B && nv + admitted[nb].values.size() <= V)
                nv += admitted[nb++].values.size();
            bool const full = nb == B || nb < admitted.size();
            bool const waited = now() - admitted.front().submitted_s >= window;
            auto free_slot = std::find_if(slots.begin(), slots.end(), [](Slot const & s) { return !s.busy; });
            if ((full || waited || stop) && free_slot != slots.end())
            {
                double const launch_start = now();
                Slot & s = *free_slot;
                std::vector<Batched> & reqs = batch_of[static_cast<size_t>(free_slot - slots.begin())];
                s.h_offs[0] = 0;
                for (uint64_t i = 0; i < nb; ++i)
                {
                    Request & r = admitted[i];
                    std::memcpy(s.h_vals + s.h_offs[i], r.values.data(), r.values.size() * 8u);
                    s.h_offs[i + 1] = s.h_offs[i] + r.values.size();
                    s.h_thr[i] = r.threshold;
                    reqs.push_back({r.ticket, r.submitted_s, r.bytes});
                }
                admitted.erase(admitted.begin(), admitted.begin() + static_cast<int64_t>(nb));
                if (int const rc = slot_launch(s, hibf, nb, nv, reinterpret_cast<char const *>(s.h_vals), 8u, R, false))
                {
                    fail(rc);
                    return;
                }
                s.busy = true;
                std::lock_guard<std::mutex> lock(m);
                ++batches;
                longest_launch_s = std::max(longest_launch_s, now() - launch_start);
            }
        }
        bool const idle = std::none_of(slots.begin(), slots.end(), [](Slot const & s) { return s.busy; });
        if (stop && admitted.empty() && idle)
            return;
        // Batches in flight: poll their events again at once.  Nothing in flight and nothing admitted:
        // sleep until a submission (or stop).  Requests waiting for their window: sleep until it ends,
        // unless a submission comes first.
        if (!idle)
        {
            std::this_thread::yield();
            continue;
        }
        std::unique_lock<std::mutex> lock(m);
        if (admitted.empty())
        {
            cv_in.wait_for(lock, std::chrono::milliseconds(2), [&] { return !incoming.empty() || stopping; });
            pass_start = now(); // an idle wait is not a stall
        }
        else
        {
            double const left = window - (now() - admitted.front().submitted_s);
            if (left > 0.0)
                cv_in.wait_for(lock, std::chrono::duration<double>(std::min(left, 1e-3)),
                               [&] { return !incoming.empty() || stopping; });
        }
    }
}

extern "C" {

int hibf_gpu_server_create(hibf_gpu_hibf hibf, hibf_gpu_serve_params const * params, hibf_gpu_server * out)
{
    if (!hibf || !params || !out)
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    hibf_gpu_serve_params const p = *params;
    if (p.streams == 0 || p.max_batch == 0 || p.results_per_query == 0 || p.max_batch_values == 0)
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT,
                            "streams, max_batch, max_batch_values and results_per_query must be > 0");
    if (p.streams > 64)
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "at most 64 streams");
    if (!(p.window_us >= 0.0) || p.window_us > 60e6)
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "window_us must be in [0, 6e7]");
    auto * s = new (std::nothrow) hibf_gpu_server_s{};
    if (!s)
        return serve_reject(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    s->hibf = hibf;
    s->p = p;
    if (hipGetDevice(&s->device) != hipSuccess)
    {
        (void)hipGetLastError();
        delete s;
        return serve_fail(hipErrorNoDevice, "get device");
    }
    s->slots.resize(p.streams);
    s->batch_of.resize(p.streams);
    for (Slot & slot : s->slots)
        if (int const rc = slot_alloc(slot, hibf, p.max_batch, p.max_batch_values, p.results_per_query, 8u, false,
                                      true))
        {
            for (Slot & f : s->slots)
                slot_free(f);
            delete s;
            return rc;
        }
    try
    {
        s->loop = std::thread([s] { s->run(); });
    }
    catch (std::exception const &)
    {
        for (Slot & f : s->slots)
            slot_free(f);
        delete s;
        return serve_reject(HIBF_GPU_ERR_OUT_OF_MEMORY, "could not start the loop's thread");
    }
    *out = s;
    return HIBF_GPU_OK;
}

int hibf_gpu_server_submit(hibf_gpu_server s, uint64_t const * h_values, uint64_t n, uint16_t threshold,
                           uint64_t * ticket)
{
    if (!s || !ticket || (!h_values && n))
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    if (n > s->p.max_batch_values)
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "request larger than max_batch_values");
    if (s->bytes_of(n) > s->p.budget_bytes)
        return serve_reject(HIBF_GPU_ERR_BUDGET, "request needs " + std::to_string(s->bytes_of(n)) +
                                                     " bytes of device memory, more than the budget");
    Request r;
    try
    {
        r.values.assign(h_values, h_values + n);
    }
    catch (std::exception const &)
    {
        return serve_reject(HIBF_GPU_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    r.threshold = threshold;
    r.bytes = s->bytes_of(n);
    {
        std::lock_guard<std::mutex> lock(s->m);
        if (s->error != HIBF_GPU_OK)
            return serve_reject(s->error, "the loop failed: " + s->error_text);
        if (s->stopping)
            return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "the server is stopping");
        r.ticket = s->next_ticket++;
        r.submitted_s = s->now();
        if (s->first_submit_s < 0.0)
            s->first_submit_s = r.submitted_s;
        *ticket = r.ticket;
        ++s->outstanding;
        s->incoming.push_back(std::move(r));
    }
    s->cv_in.notify_one();
    return HIBF_GPU_OK;
}

int hibf_gpu_server_poll(hibf_gpu_server s, hibf_gpu_server_result * results, uint64_t * h_ids, uint64_t max_results,
                         double wait_s, uint64_t * n_results)
{
    if (!s || !n_results || (max_results && (!results || !h_ids)))
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    *n_results = 0;
    uint32_t const R = s->p.results_per_query;
    std::unique_lock<std::mutex> lock(s->m);
    if (s->finished.empty() && wait_s > 0.0 && s->error == HIBF_GPU_OK)
        s->cv_out.wait_for(lock, std::chrono::duration<double>(std::min(wait_s, 3600.0)),
                           [&] { return !s->finished.empty() || s->error != HIBF_GPU_OK; });
    uint64_t k = 0;
    for (; k < max_results && !s->finished.empty(); ++k)
    {
        Finished & f = s->finished.front();
        results[k] = f.result;
        std::copy(f.ids.begin(), f.ids.end(), h_ids + k * R);
        std::fill(h_ids + k * R + f.ids.size(), h_ids + (k + 1) * R, uint64_t{0});
        s->finished.pop_front();
    }
    *n_results = k;
    if (k == 0 && s->error != HIBF_GPU_OK)
        return serve_reject(s->error, "the loop failed: " + s->error_text);
    return HIBF_GPU_OK;
}

int hibf_gpu_server_drain(hibf_gpu_server s, double wait_s)
{
    if (!s)
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    s->cv_in.notify_one();
    std::unique_lock<std::mutex> lock(s->m);
    bool const done = s->cv_out.wait_for(lock, std::chrono::duration<double>(std::min(std::max(wait_s, 0.0), 3600.0)),
                                         [&] { return s->outstanding == 0 || s->error != HIBF_GPU_OK; });
    if (s->error != HIBF_GPU_OK)
        return serve_reject(s->error, "the loop failed: " + s->error_text);
    if (!done)
        return serve_reject(HIBF_GPU_ERR_CAPACITY, std::to_string(s->outstanding) + " requests still outstanding");
    return HIBF_GPU_OK;
}

int hibf_gpu_server_stats(hibf_gpu_server s, hibf_gpu_serve_stats * stats)
{
    if (!s || !stats)
        return serve_reject(HIBF_GPU_ERR_INVALID_ARGUMENT, "null argument");
    uint64_t lookups = 0, row_bytes = 0;
    for (Slot & slot : s->slots)
    {
        uint64_t v = 0, lk = 0, rb = 0;
        if (slot.agent && hibf_gpu_agent_stats(slot.agent, &v, &lk, &rb, 0) == HIBF_GPU_OK)
        {
            lookups += lk;
            row_bytes += rb;
        }
    }
    std::lock_guard<std::mutex> lock(s->m);
    *stats = hibf_gpu_serve_stats{};
    stats->seconds = s->first_submit_s < 0.0 ? 0.0 : s->last_done_s - s->first_submit_s;
    stats->batches = s->batches;
    stats->lookups = lookups;
    stats->row_bytes = row_bytes;
    stats->max_in_flight_bytes = s->max_in_flight;
    stats->overflowed = s->overflowed;
    stats->longest_pass_s = s->longest_pass_s;
    stats->longest_launch_s = s->longest_launch_s;
    return HIBF_GPU_OK;
}

int hibf_gpu_server_destroy(hibf_gpu_server s)
{
    if (!s)
        return HIBF_GPU_OK;
    {
        std::lock_guard<std::mutex> lock(s->m);
        s->stopping = true;
    }
    s->cv_in.notify_all();
    if (s->loop.joinable())
        s->loop.join();
    for (Slot & slot : s->slots)
        slot_free(slot);
    int const rc = s->error;
    delete s;
    return rc;
}

} // extern "C"
