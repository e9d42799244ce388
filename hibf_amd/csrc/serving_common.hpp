// Pieces shared by the serving loops (serving.hip: requests replayed from a
// trace; server.hip: requests submitted while the loop runs): a slot is one
// stream with its pinned staging and device buffers and an agent, and one
// batch of requests at a time.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "hibf_gpu.h"
#include "kernels.hpp"

namespace hibf_amd
{
int set_error(int status, std::string msg);
// hibf_gpu_hibf_membership_launch whose query q writes its first min(n,
// per_query_cap) ids to d_ids[q * per_query_cap ..) and its full count n to
// d_hit_count[q] (capi.hip).
int membership_launch_capped(hibf_gpu_hibf h,
                             hibf_gpu_agent a,
                             uint64_t const * d_values,
                             uint64_t const * d_offsets,
                             uint64_t nq,
                             uint16_t const * d_thresholds,
                             uint64_t * d_ids,
                             uint32_t per_query_cap,
                             uint64_t * d_hit_start,
                             uint64_t * d_hit_count,
                             hipStream_t stream);
}

namespace hibf_amd::serving
{

// Status of a failed HIP call of the loop, with its detail for hibf_gpu_last_error.
inline int serve_fail(hipError_t e, char const * what)
{
    int const status = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? HIBF_GPU_ERR_OUT_OF_MEMORY
                                                                                    : HIBF_GPU_ERR_DEVICE;
    return set_error(status, std::string("serving loop: ") + what + ": " + hipGetErrorString(e));
}

inline int serve_reject(int status, std::string msg)
{
    return set_error(status, "serving loop: " + std::move(msg));
}

struct Slot
{
    hipStream_t stream{};
    hipEvent_t done{};
    hibf_gpu_agent agent{};
    uint64_t * h_vals{}; // pinned staging
    uint64_t * h_offs{};
    uint64_t * h_hash_offs{}; // reads in: k-mer offsets of the batch, computed on the host
    uint16_t * h_thr{};
    uint64_t * h_ids{};
    uint64_t * h_hs{};
    uint64_t * h_hc{};
    uint64_t * d_vals{};
    uint64_t * d_offs{};
    uint16_t * d_thr{};
    uint64_t * d_ids{};
    uint64_t * d_hs{};
    uint64_t * d_hc{};
    // reads in: the batch's k-mer or minimiser hashes and their offsets
    uint64_t * d_hashes{};
    uint64_t * d_hash_offs{};
    void * d_scan_tmp{}; // minimisers: the offsets scan's temporary storage
    size_t scan_bytes{};
    uint64_t q0{}, q1{};
    uint64_t bytes{};
    bool busy{false};
};

inline double now_s(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// True when both ends of [p, p + bytes) are pinned host memory known to HIP.
inline bool host_pinned(void const * p, uint64_t bytes)
{
    if (!p || !bytes)
        return false;
    for (char const * q : {static_cast<char const *>(p), static_cast<char const *>(p) + bytes - 1})
    {
        hipPointerAttribute_t a{};
        if (hipPointerGetAttributes(&a, q) != hipSuccess)
        {
            (void)hipGetLastError();
            return false;
        }
        if (a.type != hipMemoryTypeHost)
            return false;
    }
    return true;
}

// memcpy split over up to `threads` threads for copies of 8 MiB and more.
inline void copy_host(void * dst, void const * src, uint64_t bytes, unsigned threads)
{
    if (bytes < (8ull << 20) || threads <= 1)
    {
        std::memcpy(dst, src, bytes);
        return;
    }
    uint64_t const chunk = ((bytes + threads - 1) / threads + 4095u) / 4096u * 4096u;
    std::vector<std::thread> pool;
    for (uint64_t off = chunk; off < bytes; off += chunk)
        pool.emplace_back([=] { std::memcpy(static_cast<char *>(dst) + off, static_cast<char const *>(src) + off,
                                            std::min(chunk, bytes - off)); });
    std::memcpy(dst, src, std::min(chunk, bytes));
    for (std::thread & t : pool)
        t.join();
}


// Frees everything a slot holds (waits for its stream first).
inline void slot_free(Slot & s)
{
    if (s.stream)
        (void)hipStreamSynchronize(s.stream);
    for (void * d : {static_cast<void *>(s.d_vals), static_cast<void *>(s.d_offs), static_cast<void *>(s.d_thr),
                     static_cast<void *>(s.d_ids), static_cast<void *>(s.d_hs), static_cast<void *>(s.d_hc),
                     static_cast<void *>(s.d_hashes), static_cast<void *>(s.d_hash_offs), s.d_scan_tmp})
        if (d)
            (void)hipFree(d);
    for (void * hp : {static_cast<void *>(s.h_vals), static_cast<void *>(s.h_offs), static_cast<void *>(s.h_hash_offs),
                      static_cast<void *>(s.h_thr), static_cast<void *>(s.h_ids), static_cast<void *>(s.h_hs),
                      static_cast<void *>(s.h_hc)})
        if (hp)
            (void)hipHostFree(hp);
    if (s.done)
        (void)hipEventDestroy(s.done);
    if (s.stream)
        (void)hipStreamDestroy(s.stream);
    if (s.agent)
        hibf_gpu_agent_destroy(s.agent);
    s = Slot{};
}

// How reads are hashed: k-mers of length k (mini == false), or minimisers of
// (shape, window_size, seed) as hibf_gpu_minimiser_hash computes them; the
// threshold is uint16_t(fraction * number of hashes).
struct ReadHashing
{
    uint32_t k{};
    bool mini{};
    uint64_t shape{};
    uint32_t window_size{};
    uint64_t seed{};
    double fraction{};
};

// Allocates a slot for batches of at most B requests and V elements of eb bytes
// with R result ids each; staging: pinned buffers for the elements (pageable input).
inline int slot_alloc(Slot & s, hibf_gpu_hibf hibf, uint64_t B, uint64_t V, uint32_t R, uint64_t eb, bool reads,
                      bool staging, bool mini = false)
{
    auto hip = [&](hipError_t e, char const * what) { return e == hipSuccess ? HIBF_GPU_OK : serve_fail(e, what); };
    int rc = HIBF_GPU_OK;
    if (!rc) rc = hip(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "stream");
    if (!rc) rc = hip(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "event");
    if (!rc && staging)
        rc = hip(hipHostMalloc(reinterpret_cast<void **>(&s.h_vals), std::max<uint64_t>(V * eb, 8u),
                               hipHostMallocDefault), "staging");
    if (!rc) rc = hip(hipHostMalloc(reinterpret_cast<void **>(&s.h_offs), (B + 1) * 8u, hipHostMallocDefault), "offsets");
    if (!rc) rc = hip(hipHostMalloc(reinterpret_cast<void **>(&s.h_thr), B * 2u, hipHostMallocDefault), "thresholds");
    if (!rc) rc = hip(hipHostMalloc(reinterpret_cast<void **>(&s.h_ids), B * R * 8u, hipHostMallocDefault), "ids");
    if (!rc) rc = hip(hipHostMalloc(reinterpret_cast<void **>(&s.h_hs), B * 8u, hipHostMallocDefault), "hit starts");
    if (!rc) rc = hip(hipHostMalloc(reinterpret_cast<void **>(&s.h_hc), B * 8u, hipHostMallocDefault), "hit counts");
    if (!rc) rc = hip(device_malloc(&s.d_vals, std::max<uint64_t>(V * eb, 8u)), "device values");
    if (!rc && reads)
    {
        rc = hip(device_malloc(&s.d_hashes, V * 8u), "device hashes");
        if (!rc) rc = hip(device_malloc(&s.d_hash_offs, (B + 1) * 8u), "device hash offsets");
        if (!rc)
            rc = hip(hipHostMalloc(reinterpret_cast<void **>(&s.h_hash_offs), (B + 1) * 8u, hipHostMallocDefault),
                     "hash offsets");
        if (!rc && mini)
            rc = hip(minimiser_scan_bytes(B, s.scan_bytes), "minimiser scan size");
        if (!rc && mini)
            rc = hip(device_malloc(&s.d_scan_tmp, std::max<size_t>(s.scan_bytes, 8u)), "minimiser scan storage");
    }
    if (!rc) rc = hip(device_malloc(&s.d_offs, (B + 1) * 8u), "device offsets");
    if (!rc) rc = hip(device_malloc(&s.d_thr, B * 2u), "device thresholds");
    if (!rc) rc = hip(device_malloc(&s.d_ids, B * R * 8u), "device ids");
    if (!rc) rc = hip(device_malloc(&s.d_hs, B * 8u), "device hit starts");
    if (!rc) rc = hip(device_malloc(&s.d_hc, B * 8u), "device hit counts");
    if (!rc) rc = hibf_gpu_agent_create(B, &s.agent);
    if (!rc) rc = hibf_gpu_hibf_membership_prepare(hibf, s.agent, B);
    if (!rc)
    {
        uint64_t v = 0, lk = 0, rb = 0;
        (void)hibf_gpu_agent_stats(s.agent, &v, &lk, &rb, 1);
    }
    return rc;
}

// Queues one batch on a slot whose h_offs (nb + 1 offsets from 0) and, for
// hash requests, h_thr are filled: the nv elements at h_src (pinned host
// memory) to the device, reads hashed there (thresholds from the lengths),
// the membership launch, the results back and the completion event.
inline int slot_launch(Slot & s, hibf_gpu_hibf hibf, uint64_t nb, uint64_t nv, char const * h_src, uint64_t eb,
                       uint32_t R, bool reads, ReadHashing const & rh = ReadHashing{})
{
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r)
    {
        if (e == hipSuccess)
            e = r;
    };
    if (nv)
        step(hipMemcpyAsync(s.d_vals, h_src, nv * eb, hipMemcpyHostToDevice, s.stream));
    step(hipMemcpyAsync(s.d_offs, s.h_offs, (nb + 1) * 8u, hipMemcpyHostToDevice, s.stream));
    uint64_t const * d_v = s.d_vals;
    uint64_t const * d_o = s.d_offs;
    if (reads && rh.mini)
    {
        // Minimiser counts are known only once the reads are hashed: the
        // count pass, the scan and the thresholds run on the device.
        step(minimiser_hash_async(reinterpret_cast<uint8_t const *>(s.d_vals), s.d_offs, nb, rh.shape,
                                  rh.window_size, rh.seed, s.d_hash_offs, s.d_hashes, s.d_scan_tmp,
                                  std::max<size_t>(s.scan_bytes, 8u), rh.fraction, s.d_thr, s.stream));
        d_v = s.d_hashes;
        d_o = s.d_hash_offs;
    }
    else if (reads)
    {
        uint32_t const k = rh.k;
        double const fraction = rh.fraction;
        // A read of length L has max(0, L - k + 1) k-mers and the threshold
        // uint16_t(fraction * k-mers) (test/tutorial/search.cpp:42): both from
        // the lengths the host already has, so the device only hashes.
        uint64_t total = 0;
        s.h_hash_offs[0] = 0;
        for (uint64_t i = 0; i < nb; ++i)
        {
            uint64_t const len = s.h_offs[i + 1] - s.h_offs[i];
            uint64_t const windows = len >= k ? len - k + 1u : 0u;
            total += windows;
            s.h_hash_offs[i + 1] = total;
            double const t = fraction * static_cast<double>(windows);
            s.h_thr[i] = t >= 65535.0 ? uint16_t{65535} : t <= 0.0 ? uint16_t{0} : static_cast<uint16_t>(t);
        }
        step(hipMemcpyAsync(s.d_hash_offs, s.h_hash_offs, (nb + 1) * 8u, hipMemcpyHostToDevice, s.stream));
        step(hipMemcpyAsync(s.d_thr, s.h_thr, nb * 2u, hipMemcpyHostToDevice, s.stream));
        step(kmer_hash_given_offsets_async(reinterpret_cast<uint8_t const *>(s.d_vals), s.d_offs, nb, k,
                                           s.d_hash_offs, s.d_hashes, s.stream));
        d_v = s.d_hashes;
        d_o = s.d_hash_offs;
    }
    else
    {
        step(hipMemcpyAsync(s.d_thr, s.h_thr, nb * 2u, hipMemcpyHostToDevice, s.stream));
    }
    if (e != hipSuccess)
    {
        (void)hipGetLastError();
        return serve_fail(e, "batch input copies");
    }
    // Each request keeps its own first R ids, whatever the others in the batch found.
    if (int const r = membership_launch_capped(hibf, s.agent, d_v, d_o, nb, s.d_thr, s.d_ids, R, s.d_hs, s.d_hc,
                                               s.stream))
        return r;
    step(hipMemcpyAsync(s.h_hs, s.d_hs, nb * 8u, hipMemcpyDeviceToHost, s.stream));
    step(hipMemcpyAsync(s.h_hc, s.d_hc, nb * 8u, hipMemcpyDeviceToHost, s.stream));
    step(hipMemcpyAsync(s.h_ids, s.d_ids, nb * R * 8u, hipMemcpyDeviceToHost, s.stream));
    step(hipEventRecord(s.done, s.stream));
    if (e != hipSuccess)
    {
        (void)hipGetLastError();
        return serve_fail(e, "batch result copies");
    }
    return HIBF_GPU_OK;
}

// Ids of request i of a finished batch: how many of its ids, up to R, are in
// s.h_ids from s.h_hs[i] (= i * R).
inline uint64_t slot_kept_ids(Slot const & s, uint64_t i, uint64_t nb, uint32_t R)
{
    uint64_t const n = s.h_hc[i];
    uint64_t const cap = nb * R;
    uint64_t const keep = std::min<uint64_t>(n, R);
    return s.h_hs[i] > cap || keep > cap - s.h_hs[i] ? 0u : keep;
}

} // namespace hibf_amd::serving
