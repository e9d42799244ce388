// Union estimates of contiguous runs of user bins from HyperLogLog sketches
// resident on the device (the layout DP's union_estimates,
// src/sketch/toolbox.cpp:35-70).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <unordered_map>
#include <vector>

#include "layout.hpp"

namespace hibf_amd
{

// Widest sketch the device union estimates take: 2^12 registers (a thread
// keeps its lane's 512 registers in 128 VGPRs).
constexpr uint32_t kUnionMaxBits = 12u;

// est(k, j) for 0 <= k < j < np: the estimate of the union of the sketches
// regs[pos[k]] .. regs[pos[j]], truncated to uint64 (the reference's
// static_cast<uint64_t>(merge_and_estimate)).  Column j's j values start at
// j (j - 1) / 2 in `out` (np (np - 1) / 2 values).  lanes: the sketches in
// lane-major order (hll_lane_major).  Asynchronous on s.
hipError_t hll_union_triangle(uint8_t const * d_lanes,
                              uint32_t bits,
                              uint64_t const * d_pos,
                              uint64_t np,
                              uint64_t * d_out,
                              hipStream_t s);

// lanes[u][L][q] = regs[u][8 q + 7 - L]: lane L of the reference's eight float
// lanes sums exactly these registers, in this order.
hipError_t hll_lane_major(uint8_t const * d_regs, uint64_t n, uint32_t bits, uint8_t * d_lanes, hipStream_t s);

// host::UnionEstimates over device sketches: start() computes the level's whole
// triangle of estimates on the device, prefix() and column() copy it back in
// windows of columns.
class DeviceUnionEstimates final : public host::UnionEstimates
{
public:
    ~DeviceUnionEstimates() override;
    // d_registers: n sketches of 2^bits registers (device; copied, not kept).
    static int create(uint8_t const * d_registers, uint64_t n, uint32_t bits, std::unique_ptr<DeviceUnionEstimates> & out);

    bool start(std::vector<uint64_t> const & positions) override;
    bool prefix(uint64_t * prefix) override;
    uint64_t const * column(uint64_t j) override;

    double kernel_seconds{}; // triangle kernels (HIP events)
    uint64_t estimates{};    // estimates computed

private:
    DeviceUnionEstimates() = default;
    bool fetch(uint64_t j);

    uint32_t bits_{};
    uint64_t n_{};
    uint64_t np_{};
    uint8_t * d_lanes_{};
    uint64_t * d_pos_{};
    uint64_t pos_cap_{};
    uint64_t * d_tri_{};
    uint64_t tri_cap_{};
    uint64_t * h_window_{};   // pinned
    uint64_t window_cap_{};   // values
    uint64_t window_j0_{};    // columns [window_j0_, window_j1_) are on the host
    uint64_t window_j1_{};
    hipStream_t s_{};
    hipEvent_t e0_{}, e1_{};
};

// The clustering's merged-sketch estimates (toolbox::cluster_bins) on the
// device: sketches are uploaded (lane-major) the first time they are seen in a
// clustering and kept in device slots until reset(); each call then uploads
// the list of slots, estimates every union on the device (8 threads per pair,
// the reference's float lanes) and copies the doubles back.  Errors throw
// std::runtime_error after setting the library's error message.
class DevicePairUnions final : public seqan::hibf::sketch::toolbox::detail::union_source
{
public:
    ~DevicePairUnions() override;
    static int create(uint32_t bits, std::unique_ptr<DevicePairUnions> & out);

    void reset() override;
    void unions(seqan::hibf::sketch::hyperloglog const & a,
                std::vector<seqan::hibf::sketch::hyperloglog const *> const & others,
                double * out) override;
    // One launch for every pair of the leaves (the clustering's first distances).
    void all_pairs(std::vector<seqan::hibf::sketch::hyperloglog const *> const & leaves, double * out) override;

    uint64_t calls{};
    uint64_t pairs{};

private:
    DevicePairUnions() = default;
    uint32_t slot_of(seqan::hibf::sketch::hyperloglog const * h);
    void upload_pending();
    void ensure(uint64_t pairs_needed, uint64_t slots_needed);
    void check(hipError_t e, char const * what);

    uint32_t bits_{};
    std::unordered_map<seqan::hibf::sketch::hyperloglog const *, uint32_t> slots_;
    std::vector<seqan::hibf::sketch::hyperloglog const *> pending_; // new this call, in slot order
    uint64_t used_{};      // slots in use
    uint64_t uploaded_{};  // slots already on the device
    uint8_t * d_slots_{};
    uint64_t slot_cap_{};
    uint8_t * h_stage_{};  // pinned: lane-major copies of new sketches
    uint64_t stage_cap_{}; // sketches
    uint32_t * d_idx_{};
    double * d_out_{};
    uint32_t * h_idx_{};   // pinned
    double * h_out_{};     // pinned
    uint64_t pair_cap_{};
    double const * d_table_{};
    double norm_{};
    hipStream_t s_{};
};

} // namespace hibf_amd
