"""MI355X-native HIBF query path.

libhibf_gpu.so (hand-written HIP kernels for gfx950 behind the C-ABI of
include/hibf_gpu.h) is the compute path; this package binds it for Python
callers.  `hibf_amd.device` holds the device-resident IBF/HIBF handles.
"""
from .native import HibfGpuError, lib  # noqa: F401

__all__ = ["HibfGpuError", "lib"]
