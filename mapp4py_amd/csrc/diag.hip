// diag.hip -- plain HBM copy kernel of known size. It calibrates the HBM
// traffic counters (FETCH_SIZE / WRITE_SIZE read for a copy of B bytes) and
// gives a measured device-bandwidth figure next to the 8 TB/s peak.
#include "../../include/mapp_amd.h"
#include "common.hpp"

namespace mapp {
namespace {
__global__ __launch_bounds__(256) void k_copy16(size_t n16, const float4* __restrict__ a, float4* __restrict__ b)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) b[i] = a[i];
}
}  // namespace
}  // namespace mapp

extern "C" mapp_status mapp_diag_copy(size_t bytes, int reps, double* gbps)
{
    try {
        MAPP_REQUIRE(bytes >= 16 && reps >= 1 && gbps, mapp::ERR_ARG, "invalid copy calibration arguments");
        const size_t n16 = bytes / 16;
        mapp::DevBuf<float4> a, b;
        a.reserve(n16);
        b.reserve(n16);
        MAPP_HIP(hipMemset(a.get(), 0, n16 * 16));
        hipEvent_t e0, e1;
        MAPP_HIP(hipEventCreate(&e0));
        MAPP_HIP(hipEventCreate(&e1));
        hipLaunchKernelGGL(mapp::k_copy16, dim3(2048), dim3(256), 0, 0, n16, a.get(), b.get());
        MAPP_HIP(hipEventRecord(e0, 0));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(mapp::k_copy16, dim3(2048), dim3(256), 0, 0, n16, a.get(), b.get());
        MAPP_HIP(hipEventRecord(e1, 0));
        MAPP_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        MAPP_HIP(hipEventElapsedTime(&ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *gbps = 2.0 * (double)n16 * 16.0 * reps / (ms * 1e-3) / 1e9;
        return MAPP_OK;
    } catch (const mapp::Error& e) {
        mapp::set_last_error(e.what());
        return e.code;
    }
}
