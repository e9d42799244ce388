// The device side of the C-ABI for host-only test builds (the sanitizer run of
// the host code): every entry point that needs the GPU reports that there is
// none, so the host parsers and the build plan can be exercised on the CPU.
#include <string>

#include "hibf_gpu.h"

namespace
{
thread_local std::string g_error;
}

namespace hibf_amd
{
int set_error(int status, std::string msg)
{
    g_error = std::move(msg);
    return status;
}
} // namespace hibf_amd

extern "C" {
const char * hibf_gpu_last_error(void) { return g_error.c_str(); }
int hibf_gpu_ibf_create(uint64_t, uint64_t, uint32_t, hibf_gpu_ibf *) { return HIBF_GPU_ERR_NO_DEVICE; }
int hibf_gpu_ibf_destroy(hibf_gpu_ibf) { return HIBF_GPU_OK; }
int hibf_gpu_ibf_get_geometry(hibf_gpu_ibf, hibf_gpu_ibf_geometry *) { return HIBF_GPU_ERR_NO_DEVICE; }
int hibf_gpu_ibf_device_words(hibf_gpu_ibf, uint64_t **) { return HIBF_GPU_ERR_NO_DEVICE; }
int hibf_gpu_ibf_occupancy(hibf_gpu_ibf, uint64_t *, int *) { return HIBF_GPU_ERR_NO_DEVICE; }
int hibf_gpu_ibf_set_occupancy(hibf_gpu_ibf, uint64_t const *, int) { return HIBF_GPU_ERR_NO_DEVICE; }
int hibf_gpu_hibf_create(uint64_t, hibf_gpu_ibf_geometry const *, uint64_t const *, uint64_t const *,
                         uint64_t const *, uint64_t const *, uint64_t const *, uint64_t, hibf_gpu_hibf *)
{
    return HIBF_GPU_ERR_NO_DEVICE;
}
int hibf_gpu_hibf_destroy(hibf_gpu_hibf) { return HIBF_GPU_OK; }
int hibf_gpu_hibf_info(hibf_gpu_hibf, uint64_t *, uint64_t *, uint64_t *, uint64_t *, uint64_t *)
{
    return HIBF_GPU_ERR_NO_DEVICE;
}
int hibf_gpu_hibf_export_tables(hibf_gpu_hibf, hibf_gpu_ibf_geometry *, uint64_t *, uint64_t *, uint64_t *,
                                uint64_t *)
{
    return HIBF_GPU_ERR_NO_DEVICE;
}
int hibf_gpu_hibf_device_arena(hibf_gpu_hibf, uint64_t **) { return HIBF_GPU_ERR_NO_DEVICE; }
}
