// Internal: growable device buffer over the C-ABI's memory helpers, so the
// seqan::hibf headers need no HIP headers.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#include <hibf_gpu.h>

namespace seqan::hibf::detail
{

[[noreturn]] inline void throw_status(int status)
{
    std::string msg = hibf_gpu_last_error();
    switch (status)
    {
    case HIBF_GPU_ERR_BIN_COUNT:
    case HIBF_GPU_ERR_HASH_COUNT:
    case HIBF_GPU_ERR_BIN_SIZE: throw std::logic_error{msg};
    case HIBF_GPU_ERR_INVALID_ARGUMENT:
    case HIBF_GPU_ERR_CONFIG: throw std::invalid_argument{msg};
    default: throw std::runtime_error{std::string{hibf_gpu_status_string(status)} + ": " + msg};
    }
}

inline void check(int status)
{
    if (status != HIBF_GPU_OK)
        throw_status(status);
}

class device_buffer
{
public:
    device_buffer() = default;
    // A copy is a fresh, empty buffer: buffers hold per-call scratch only.
    device_buffer(device_buffer const &) noexcept {}
    device_buffer & operator=(device_buffer const &) noexcept { return *this; }
    device_buffer(device_buffer && o) noexcept : ptr_{o.ptr_}, bytes_{o.bytes_}
    {
        o.ptr_ = nullptr;
        o.bytes_ = 0;
    }
    device_buffer & operator=(device_buffer && o) noexcept
    {
        std::swap(ptr_, o.ptr_);
        std::swap(bytes_, o.bytes_);
        return *this;
    }
    ~device_buffer() { hibf_gpu_free(ptr_); }

    void * reserve(size_t bytes)
    {
        if (bytes > bytes_)
        {
            hibf_gpu_free(ptr_);
            ptr_ = nullptr;
            bytes_ = 0;
            check(hibf_gpu_malloc(&ptr_, bytes));
            bytes_ = bytes;
        }
        return ptr_;
    }
    template <typename T>
    T * as() const noexcept
    {
        return static_cast<T *>(ptr_);
    }
    void upload(void const * src, size_t bytes)
    {
        reserve(bytes);
        check(hibf_gpu_memcpy(ptr_, src, bytes, 0));
    }
    void download(void * dst, size_t bytes) const { check(hibf_gpu_memcpy(dst, ptr_, bytes, 1)); }

private:
    void * ptr_{nullptr};
    size_t bytes_{0};
};

} // namespace seqan::hibf::detail
