// hll_host.cpp with the device check (compute_sketches) compiled in; built
// with the other C++ test programs and linked against libhibf_gpu.so.
#define WITH_DEVICE 1
#include "hll_host.cpp"
