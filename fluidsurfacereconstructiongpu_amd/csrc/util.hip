// Small device utilities shared by the host code.
#include "common.hpp"

namespace fsr {

namespace {

// One lane reads the constant-rate (100 MHz) device clock and stores it.
__global__ void k_timestamp(uint64_t* out) {
    if (threadIdx.x == 0) *out = __builtin_amdgcn_s_memrealtime();
}

}  // namespace

fsr_status device_timestamp(uint64_t* out, hipStream_t s) {
    hipLaunchKernelGGL(k_timestamp, dim3(1), dim3(64), 0, s, out);
    FSR_LAUNCH_CHECK("k_timestamp");
    return FSR_OK;
}

}  // namespace fsr
