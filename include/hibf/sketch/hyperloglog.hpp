// seqan::hibf::sketch::hyperloglog on the host, with the reference's members
// and rules (include/hibf/sketch/hyperloglog.hpp, src/sketch/hyperloglog.cpp):
// 2^bits uint8 registers; add(v) hashes v as lo64(v * C) ^ hi64(v * C) with
// C = 0x9E3779B97F4A7C15, the top `bits` bits pick the register, which keeps
// the max of (leading zeros of the remaining bits, at most 64 - bits) + 1;
// estimate() sums 2^-register in eight float lanes in the reference's order
// and applies alpha m^2 / sum with the small-value correction
// m ln(m / zeros).  The device computes the same registers for many sets at
// once (hibf_gpu_hll_sketch, used by compute_sketches).
#pragma once

#include <bit>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <vector>

namespace seqan::hibf::sketch
{

class hyperloglog
{
public:
    // hyperloglog.cpp:26-46
    hyperloglog(uint8_t const num_bits = 5u) : bits{num_bits}
    {
        if (bits < 5u || bits > 32u)
            throw std::invalid_argument("[HyperLogLog] bit width must be in the range [5,32].");
        size = uint64_t{1} << bits;
        data.assign(size, 0u);
        double const alpha = size == 32u ? 0.697 : size == 64u ? 0.709 : 0.7213 / (1.0 + 1.079 / size);
        normalization_factor = alpha * size * size;
        rank_mask = (uint64_t{1} << bits) - 1u;
    }

    // hyperloglog.cpp:56-64
    void add(uint64_t const value)
    {
        unsigned __int128 const p = static_cast<unsigned __int128>(value) * 0x9E3779B97F4A7C15ULL;
        uint64_t const hash = static_cast<uint64_t>(p) ^ static_cast<uint64_t>(p >> 64);
        uint64_t const index = hash >> (64u - bits);
        uint8_t const rank = static_cast<uint8_t>(std::countl_zero((hash << bits) | rank_mask) + 1);
        data[index] = rank > data[index] ? rank : data[index];
    }

    // hyperloglog.cpp:66-129
    double estimate() const
    {
        float lane[8]{};
        for (uint64_t block = 0; block < size / 32u; ++block)
        {
            uint8_t const * const cur = data.data() + 32u * block;
            for (size_t j = 0; j <= 24u; j += 8u)
                for (size_t l = 0; l < 8u; ++l) // lane l holds register j + 7 - l (_mm256_set_ps order)
                    lane[l] += 1.0f / static_cast<float>(uint64_t{1} << cur[j + 7u - l]);
        }
        float sum{};
        for (float const x : lane)
            sum += x;
        double estimate = normalization_factor / static_cast<double>(sum);
        if (estimate <= 2.5 * size)
        {
            uint32_t zeros{};
            for (uint8_t const r : data)
                zeros += r == 0u;
            if (zeros != 0u)
                estimate = size * std::log(static_cast<double>(size) / zeros);
        }
        return estimate;
    }

    // hyperloglog.cpp:136-149
    void merge(hyperloglog const & other)
    {
        for (uint64_t i = 0; i < size; ++i)
            data[i] = other.data[i] > data[i] ? other.data[i] : data[i];
    }

    double merge_and_estimate(hyperloglog const & other)
    {
        merge(other);
        return estimate();
    }

    void reset() { data.assign(size, 0u); }

    uint64_t data_size() const { return size; }

    // hyperloglog.cpp:163-199: one byte with the bit width, then the registers.
    void store(std::ostream & os) const
    {
        os.write(reinterpret_cast<char const *>(&bits), 1);
        os.write(reinterpret_cast<char const *>(data.data()), static_cast<std::streamsize>(size));
        os.flush();
        if (os.fail())
            throw std::runtime_error("[HyperLogLog] Failed to store a HyperLogLog sketch to a file.");
    }

    void load(std::istream & is)
    {
        uint8_t restore_bits{};
        is.read(reinterpret_cast<char *>(&restore_bits), 1);
        hyperloglog restored;
        try
        {
            restored = hyperloglog{restore_bits};
        }
        catch (std::invalid_argument const &)
        {
            throw std::runtime_error("[HyperLogLog] Failed to load a HyperLogLog sketch from a file: Invalid bit_width.");
        }
        is.read(reinterpret_cast<char *>(restored.data.data()), static_cast<std::streamsize>(restored.size));
        if (is.fail())
            throw std::runtime_error("[HyperLogLog] Failed to load a HyperLogLog sketch from a file: I/O error.");
        *this = std::move(restored);
    }

    // Not in the reference: the registers, and a sketch from registers computed elsewhere
    // (the device sketches of compute_sketches).
    std::vector<uint8_t> const & registers() const noexcept { return data; }
    static hyperloglog from_registers(uint8_t const num_bits, uint8_t const * regs)
    {
        hyperloglog h{num_bits};
        h.data.assign(regs, regs + h.size);
        return h;
    }

    bool operator==(hyperloglog const &) const = default;

private:
    uint8_t bits{};
    uint64_t size{};
    uint64_t rank_mask{};
    double normalization_factor{};
    std::vector<uint8_t> data{};
};

} // namespace seqan::hibf::sketch
